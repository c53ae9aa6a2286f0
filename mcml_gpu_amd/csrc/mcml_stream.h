/* Internal to libmcml: result arrays reported as they reach the host during a batch run.
 *
 * The executable (io.cpp mcml_main) writes the pathlength files of each (set, SDS) group as
 * soon as that group's rows have been read back, on a thread of its own, while the rest of
 * the rows and the absorbance grids are still crossing PCIe.  The reference writes every
 * file after the whole simulation (transport.cu:156-251, io.cpp:281-469). */
#ifndef MCML_STREAM_H
#define MCML_STREAM_H

#include <stdint.h>

#include "../../include/mcml.h"

struct mcml_readback_sink {
    /* Rows [0, rows) of res->pathlength are on the host, and res->pl_offset is complete.
     * Called on the thread running the batch, with growing `rows`, once per read-back chunk. */
    void (*rows_landed)(void* user, const mcml_result* res, uint64_t rows);
    /* Called before the batch frees the result on an error, once rows_landed has been
     * called; returns when the caller no longer reads the result. */
    void (*release)(void* user);
    void* user;
};

/* mcml_run_batch, reporting the pathlength rows to `sink` (may be null) as they land. */
int mcml_run_batch_sink(mcml_ctx* c, const mcml_sim* sims, int nsets, const mcml_opts* opts, uint64_t seed,
                        mcml_result* out, const mcml_readback_sink* sink);

#endif
