/* MCML_GPU drop-in executable: the reference's command line (main.cu:17-145) over libmcml. */
#include "../../include/mcml.h"

int main(int argc, char** argv) { return mcml_main(argc, argv); }
