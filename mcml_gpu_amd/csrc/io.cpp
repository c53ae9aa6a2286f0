/*
 * io.cpp - the operator surface of the reference program, above the C ABI:
 *   - sim_set.json + input.txt reading and validation (read_mua_mus, io.cpp:35-256),
 *     with the reference's messages; a small JSON reader replaces the vendored
 *     nlohmann/json (only parsing is needed);
 *   - the output writers, byte-compatible with io.cpp:11-33, 258-469 (std::ofstream with
 *     default formatting, append mode for text, overwrite for binary);
 *   - the MCML_GPU command line (main.cu:17-180) as mcml_main.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mcml.h"
#include "mcml_stream.h"

namespace {

constexpr float kPI = 3.141592654f; /* header.h:40 */
constexpr const char* kVersion = "3.02";
constexpr const char* kUpdated = "2020/06/13";

/* ------------------------------------------------------------------ tiny JSON reader */
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const std::string& k) const
    {
        if (kind != OBJ) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;

    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool fail(const char* m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char* s)
    {
        size_t n = strlen(s);
        if ((size_t)(end - p) < n || strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string& out)
    {
        if (p >= end || *p != '"') return fail("expected string");
        p++;
        out.clear();
        while (p < end && *p != '"') {
            if (*p == '\\') {
                p++;
                if (p >= end) return fail("bad escape");
                char c = *p++;
                switch (c) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u':
                    if (end - p < 4) return fail("bad \\u escape");
                    out += '?';
                    p += 4;
                    break;
                default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(JVal& v, int depth)
    {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        char c = *p;
        if (c == '{') {
            p++;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') {
                p++;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                p++;
                JVal child;
                if (!value(child, depth + 1)) return false;
                v.obj.emplace_back(k, std::move(child));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            p++;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') {
                p++;
                return true;
            }
            for (;;) {
                JVal child;
                if (!value(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::STR;
            return string(v.str);
        }
        if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JVal::BOOL;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.kind = JVal::NUL;
            return true;
        }
        char* q = nullptr;
        double d = strtod(p, &q);
        if (q == p) return fail("invalid value");
        p = q;
        v.kind = JVal::NUM;
        v.num = d;
        return true;
    }
};

bool parse_json(const std::string& text, JVal& root, std::string& err)
{
    JParser jp{text.data(), text.data() + text.size(), {}};
    if (!jp.value(root, 0)) {
        err = jp.err;
        return false;
    }
    jp.ws();
    if (jp.p != jp.end) {
        err = "trailing characters";
        return false;
    }
    return true;
}

std::string jtext(const JVal* v)
{
    if (!v) return "null";
    std::ostringstream os;
    switch (v->kind) {
    case JVal::NUM: os << v->num; break;
    case JVal::BOOL: os << (v->b ? "true" : "false"); break;
    case JVal::STR: os << '"' << v->str << '"'; break;
    case JVal::NUL: os << "null"; break;
    default: os << "<composite>"; break;
    }
    return os.str();
}

bool jnum(const JVal* v, double& out)
{
    if (!v) return false;
    if (v->kind == JVal::NUM) {
        out = v->num;
        return true;
    }
    if (v->kind == JVal::BOOL) {
        out = v->b ? 1.0 : 0.0;
        return true;
    }
    return false;
}

void copy_msg(char* msg, size_t msglen, const std::string& s)
{
    if (msg && msglen) {
        strncpy(msg, s.c_str(), msglen - 1);
        msg[msglen - 1] = '\0';
    }
}

std::string filename(const char* outdir, const std::string& name)
{
    if (!outdir || !*outdir) return name;
    std::string d(outdir);
    if (d.back() != '/') d += '/';
    return d + name;
}

/* Appending text writer with the bytes of `std::ostream << double` at the default precision
 * (printf "%g", which is what the stream's num_put formats with), gathered in a 1 MB buffer
 * instead of flushing every line.  Large text outputs (500 k A_rz values per SDS at the
 * default grid, 1e5+ pathlength records) are written several times faster. */
class TextFile {
  public:
    explicit TextFile(const std::string& path) : f_(fopen(path.c_str(), "ab")) { buf_.reserve(kCap + 64); }
    ~TextFile() { close(); }
    bool ok() const { return f_ != nullptr && !bad_; }
    void num(double v)
    {
        char t[40];
        const int n = snprintf(t, sizeof(t), "%g", v);
        buf_.append(t, (size_t)n);
        if (buf_.size() >= kCap) flush();
    }
    void put(char c) { buf_.push_back(c); }
    /* append text formatted elsewhere (already in this file's format) */
    void raw(const std::string& s)
    {
        flush();
        if (f_ && !s.empty() && fwrite(s.data(), 1, s.size(), f_) != s.size()) bad_ = true;
    }
    /* flush and close; false if anything failed */
    bool close()
    {
        if (!f_) return false;
        flush();
        if (fclose(f_) != 0) bad_ = true;
        f_ = nullptr;
        return !bad_;
    }

  private:
    static constexpr size_t kCap = 1 << 20;
    void flush()
    {
        if (f_ && !buf_.empty() && fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) bad_ = true;
        buf_.clear();
    }
    FILE* f_;
    std::string buf_;
    bool bad_ = false;
};

void append_g(std::string& out, double v)
{
    char t[40];
    const int n = snprintf(t, sizeof(t), "%g", v);
    out.append(t, (size_t)n);
}

/* Format rows [0, n) of `per_row` values each with row(i, out) on up to 16 host threads (one
 * per 32 k values), each into its own string, and append them to f in row order. */
template <class RowFn>
void write_rows_parallel(TextFile& f, uint64_t n, uint64_t per_row, RowFn row)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const uint64_t nt = std::min<uint64_t>({(uint64_t)std::max(1u, std::min(hw, 16u)), std::max<uint64_t>(1, n * per_row / 32768),
                                            std::max<uint64_t>(1, n)});
    std::vector<std::string> parts((size_t)nt);
    auto work = [&](uint64_t t) {
        const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        for (uint64_t i = b; i < e; i++) row(i, parts[(size_t)t]);
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (uint64_t t = 0; t < nt; t++) th.emplace_back(work, t);
        for (std::thread& x : th) x.join();
    }
    for (const std::string& s : parts) f.raw(s);
}

} // namespace

/* ====================================================================== input */
extern "C" int mcml_read_input(const char* sim_set_json, const char* input_txt, mcml_sim** sims, int* nsims,
                               char* msg, size_t msglen)
{
    if (!sim_set_json || !input_txt || !sims || !nsims) return MCML_ERR_ARG;
    *sims = nullptr;
    *nsims = 0;
    std::ifstream jf(sim_set_json);
    if (!jf) {
        copy_msg(msg, msglen, std::string("cannot open ") + sim_set_json + "\n");
        return MCML_ERR_IO;
    }
    std::stringstream ss;
    ss << jf.rdbuf();
    JVal root;
    std::string jerr;
    if (!parse_json(ss.str(), root, jerr) || root.kind != JVal::OBJ) {
        copy_msg(msg, msglen, std::string("sim_set.json parse error: ") + (jerr.empty() ? "not an object" : jerr) + "\n");
        return MCML_ERR_INPUT;
    }
    std::ostringstream m;
    double v;
    const JVal* probes = root.get("probes");
    const JVal* src = probes ? probes->get("source") : nullptr;
    const JVal* dets = probes ? probes->get("detectors") : nullptr;
    double n_photons, n_sim, n_layers, n_sds, refl, upper_n, lower_n;
    if (!jnum(root.get("number_photons"), n_photons) || !jnum(root.get("number_simulation"), n_sim) ||
        !jnum(root.get("number_layers"), n_layers) || !jnum(root.get("detector_reflectance"), refl) ||
        !jnum(root.get("upper_n"), upper_n) || !jnum(root.get("lower_n"), lower_n) || !src || !dets ||
        dets->kind != JVal::ARR || !jnum(probes->get("num_SDS"), n_sds)) {
        copy_msg(msg, msglen, "sim_set.json is missing a required key\n");
        return MCML_ERR_INPUT;
    }
    /* io.cpp:52-80 */
    const JVal* so = root.get("source_probe_oblique");
    const JVal* dobl = root.get("detector_probe_oblique");
    int src_obl, det_obl;
    if (jnum(so, v) && v == 1) src_obl = 1;
    else if (jnum(so, v) && v == 0) src_obl = 0;
    else {
        m << "source_probe_oblique: " << jtext(so) << "is not 1 or 0!\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    if (jnum(dobl, v) && v == 1) det_obl = 1;
    else if (jnum(dobl, v) && v == 0) det_obl = 0;
    else {
        m << "detector_probe_oblique: " << jtext(dobl) << "is not 1 or 0!\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    if (refl < 0 || refl > 1) { /* io.cpp:82-86 */
        m << "Detector reflectance: " << refl << " is out of range !\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    const int L = (int)n_layers;
    if (L > MCML_MAX_LAYERS) { /* io.cpp:88-92 */
        m << "Number of layer is too large!\nThis program only allow number lower than " << MCML_MAX_LAYERS << " !\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    const int S = (int)n_sds;
    if (S > MCML_MAX_SDS) { /* io.cpp:97-101 */
        m << "Number of SDS is too large!\nThis program only allow number lower than " << MCML_MAX_SDS << " !\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    if ((size_t)S != dets->arr.size()) { /* io.cpp:102-105 */
        m << "Number of SDS : " << S << " not match the probe setting : " << dets->arr.size() << " !\n";
        copy_msg(msg, msglen, m.str());
        return MCML_ERR_INPUT;
    }
    const int NS = (int)n_sim;
    if (L < 1 || S < 1 || NS < 1) {
        copy_msg(msg, msglen, "number_layers, num_SDS and number_simulation must be positive\n");
        return MCML_ERR_INPUT;
    }
    mcml_sim base;
    memset(&base, 0, sizeof(base));
    base.number_photons = (uint64_t)n_photons;
    base.number_layers = L;
    base.num_sds = S;
    base.detector_reflectance = (float)refl;
    base.upper_n = (float)upper_n;
    base.lower_n = (float)lower_n;
    base.source_probe_oblique = src_obl;
    base.detector_probe_oblique = det_obl;
    double na = 0, rad = 0, ang = 0, pos = 0;
    jnum(src->get("NA"), na);
    jnum(src->get("radius"), rad);
    jnum(src->get("angle"), ang);
    base.source.pos = 0.0f;
    base.source.NA = (float)na;
    base.source.radius = (float)rad;
    base.source.angle_deg = (float)ang;
    for (int d = 0; d < S; d++) {
        const JVal& dj = dets->arr[(size_t)d];
        na = rad = ang = pos = 0;
        jnum(dj.get("pos"), pos);
        jnum(dj.get("NA"), na);
        jnum(dj.get("radius"), rad);
        jnum(dj.get("angle"), ang);
        base.detectors[d].pos = (float)pos;
        base.detectors[d].NA = (float)na;
        base.detectors[d].radius = (float)rad;
        base.detectors[d].angle_deg = (float)ang;
    }
    /* optional absorbance-grid keys (not in the reference, which compiles them in) */
    if (jnum(root.get("grid_nr"), v)) base.grid_nr = (int)v;
    if (jnum(root.get("grid_nz"), v)) base.grid_nz = (int)v;
    if (jnum(root.get("grid_dr"), v)) base.grid_dr = (float)v;
    if (jnum(root.get("grid_dz"), v)) base.grid_dz = (float)v;

    /* input.txt: whitespace-separated, 5L-1 numbers per simulation (io.cpp:135-161);
     * a failed extraction leaves 0, like operator>> */
    std::ifstream tf(input_txt);
    if (!tf) {
        copy_msg(msg, msglen, std::string("cannot open ") + input_txt + "\n");
        return MCML_ERR_IO;
    }
    std::vector<mcml_sim> out((size_t)NS, base);
    auto next = [&tf]() -> float {
        float x = 0.0f;
        if (!(tf >> x)) x = 0.0f;
        return x;
    };
    for (int i = 0; i < NS; i++) {
        mcml_sim& s = out[(size_t)i];
        int j = 0;
        for (; j < L - 1; j++) {
            s.layers[j].thickness = next();
            s.layers[j].mua = next();
            s.layers[j].mus = next();
            s.layers[j].n = next();
            s.layers[j].g = next();
        }
        s.layers[j].thickness = 0.0f;
        s.layers[j].mua = next();
        s.layers[j].mus = next();
        s.layers[j].n = next();
        s.layers[j].g = next();
        if (s.layers[j].mus <= 0) { /* io.cpp:149-160: only the last layer is checked */
            m << "Mus for simulation " << i << " layer " << j << " Error with value " << s.layers[j].mus << " !\n";
            copy_msg(msg, msglen, m.str());
            return MCML_ERR_INPUT;
        }
        if (s.layers[j].n <= 1) {
            m << "n for simulation " << i << " layer " << j << " Error with value " << s.layers[j].n << " !\n";
            copy_msg(msg, msglen, m.str());
            return MCML_ERR_INPUT;
        }
        if (s.layers[j].g > 1) {
            m << "g for simulation " << i << " layer " << j << " Error with value " << s.layers[j].g << " !\n";
            copy_msg(msg, msglen, m.str());
            return MCML_ERR_INPUT;
        }
    }
    mcml_sim* arr = (mcml_sim*)malloc(sizeof(mcml_sim) * (size_t)NS);
    if (!arr) return MCML_ERR_NOMEM;
    memcpy(arr, out.data(), sizeof(mcml_sim) * (size_t)NS);
    *sims = arr;
    *nsims = NS;
    copy_msg(msg, msglen, "");
    return MCML_OK;
}

extern "C" void mcml_free_sims(mcml_sim* sims) { free(sims); }

/* ====================================================================== writers */
/* -P: pathlength_SDS_#.bin / .txt of one (set, SDS) group (io.cpp:281-329).  Reads only
 * res->pl_offset and the group's rows of res->pathlength. */
static int write_pathlength_file(const mcml_opts* opts, const mcml_result* res, int set, int s, const char* outdir)
{
    const int S = res->num_sds, L = res->num_layers;
    const size_t g = (size_t)set * (size_t)S + (size_t)s;
    const uint64_t b = res->pl_offset[g], e = res->pl_offset[g + 1];
    const float* recs = res->pathlength + b * (uint64_t)(L + 2);
    if (opts->binary) {
        std::string fn = filename(outdir, "pathlength_SDS_" + std::to_string(s + 1) + ".bin");
        FILE* f = fopen(fn.c_str(), "wb");
        if (!f) return MCML_ERR_IO;
        size_t n = (size_t)(e - b) * (size_t)(L + 2);
        if (n && fwrite(recs, sizeof(float), n, f) != n) {
            fclose(f);
            return MCML_ERR_IO;
        }
        return fclose(f) == 0 ? MCML_OK : MCML_ERR_IO;
    }
    TextFile f(filename(outdir, "pathlength_SDS_" + std::to_string(s + 1) + ".txt"));
    if (!f.ok()) return MCML_ERR_IO;
    write_rows_parallel(f, e - b, (uint64_t)(L + 2), [&](uint64_t i, std::string& out) {
        for (int j = 0; j <= L + 1; j++) {
            append_g(out, recs[i * (L + 2) + j]);
            out.push_back('\t');
        }
        out.push_back('\n');
    });
    return f.close() ? MCML_OK : MCML_ERR_IO;
}

/* Every output of one simulation; `pathlength_files` = false leaves out the -P files (the
 * executable writes those while the batch is still being read back). */
static int write_outputs(const mcml_sim* sim, const mcml_opts* opts, const mcml_result* res, int set,
                         const char* output_txt, const char* outdir, const mcml_summary_info* info,
                         bool pathlength_files)
{
    const int S = res->num_sds, L = res->num_layers;
    const double scale1 = (double)0xFFFFFFFF * (double)sim->number_photons;
    const size_t g0 = (size_t)set * (size_t)S;
    const bool replay = opts->replay && res->replayed;
    if (replay && opts->pathlength && pathlength_files)
        for (int s = 0; s < S; s++) {
            const int rc = write_pathlength_file(opts, res, set, s, outdir);
            if (rc) return rc;
        }
    /* -AP: average_pathlength.txt (io.cpp:331-343) */
    if (replay && opts->avg_pathlength) {
        std::vector<double> avg((size_t)S * L);
        int rc = mcml_average_pathlength(sim, res, set, avg.data());
        if (rc) return rc;
        std::ofstream f(filename(outdir, "average_pathlength.txt"), std::ios::app);
        if (!f) return MCML_ERR_IO;
        for (int s = 0; s < S; s++)
            for (int l = 0; l < L; l++) f << avg[(size_t)s * L + l] << '\t';
        f << std::endl;
    }
    /* output.txt (io.cpp:11-33): the fiber build separates with ' ' */
    {
        std::ofstream f(output_txt, std::ios::app);
        if (!f) return MCML_ERR_IO;
        const double* refl = replay ? res->replay_reflectance : res->reflectance;
        for (int s = 0; s < S; s++) f << double(refl[g0 + s] / scale1) << " ";
        f << std::endl;
    }
    /* -A: A_rz_SDS_# and A0_z_SDS_# (io.cpp:383-469) */
    if (replay && opts->absorbance && res->A_rz && res->A0_z) {
        mcml_derived d;
        int rc = mcml_derive(sim, &d);
        if (rc) return rc;
        const int nr = res->grid_nr, nz = res->grid_nz;
        const double dr = d.grid_dr, dz = d.grid_dz;
        for (int s = 0; s < S; s++) {
            const uint64_t* arz = res->A_rz + (g0 + s) * (size_t)nz * (size_t)nr;
            std::string base = "A_rz_SDS_" + std::to_string(s + 1);
            if (opts->binary) {
                FILE* f = fopen(filename(outdir, base + ".bin").c_str(), "wb");
                if (!f) return MCML_ERR_IO;
                std::vector<float> row((size_t)nr);
                for (int z = 0; z < nz; z++) {
                    for (int r = 0; r < nr; r++) {
                        const double scale2 = scale1 * 2 * kPI * (r + 0.5) * dr * dr * dz;
                        row[(size_t)r] = float(double(arz[(size_t)z * nr + r] / scale2));
                    }
                    fwrite(row.data(), sizeof(float), (size_t)nr, f);
                }
                fclose(f);
            } else {
                TextFile f(filename(outdir, base + ".txt"));
                if (!f.ok()) return MCML_ERR_IO;
                write_rows_parallel(f, (uint64_t)nz, (uint64_t)nr, [&](uint64_t z, std::string& out) {
                    for (int r = 0; r < nr; r++) {
                        const double scale2 = scale1 * 2 * kPI * (r + 0.5) * dr * dr * dz;
                        append_g(out, double(arz[(size_t)z * nr + r] / scale2));
                        out.push_back('\t');
                    }
                    out.push_back('\n');
                });
                if (!f.close()) return MCML_ERR_IO;
            }
        }
        const double scale2 = scale1 * 2 * kPI * (0 + 0.5) * dr * dr * dz;
        for (int s = 0; s < S; s++) {
            const uint64_t* a0 = res->A0_z + (g0 + s) * (size_t)nz;
            std::string base = "A0_z_SDS_" + std::to_string(s + 1);
            if (opts->binary) {
                FILE* f = fopen(filename(outdir, base + ".bin").c_str(), "wb");
                if (!f) return MCML_ERR_IO;
                for (int z = 0; z < nz; z++) {
                    float t = float(double(a0[z] / scale2));
                    fwrite(&t, sizeof(float), 1, f);
                }
                fclose(f);
            } else {
                TextFile f(filename(outdir, base + ".txt"));
                if (!f.ok()) return MCML_ERR_IO;
                for (int z = 0; z < nz; z++) {
                    f.num(double(a0[z] / scale2));
                    f.put('\n');
                }
                if (!f.close()) return MCML_ERR_IO;
            }
        }
    }
    /* summary.json (io.cpp:345-381) */
    {
        std::ofstream f(filename(outdir, "summary.json"), std::ios::app);
        if (!f) return MCML_ERR_IO;
        const double t_sim = info ? info->sim_time_s : 0.0;
        const double t_rep = info ? info->replay_time_s : 0.0;
        double speed;
        f << "{\n";
        f << "\"MCML_version_K\": " << kVersion << ",\n";
        f << "\"num_photon\": " << (unsigned long long)sim->number_photons << ",\n";
        f << "\"sim_time\": " << t_sim << ",\n";
        if (replay) {
            f << "\"replay_time\": " << t_rep << ",\n";
            speed = (double)sim->number_photons / (t_sim + t_rep);
        } else {
            speed = (double)sim->number_photons / t_sim;
        }
        f << "\"sim_speed(photons/s)\": " << speed << ",\n";
        f << "\"sim_GPU\": \"" << (info && info->gpu_name ? info->gpu_name : "") << "\",\n";
        f << "\"each_photon_weight\": " << scale1 << ",\n";
        f << "\"number_layers\": " << L << ",\n";
        f << "\"detect_mode\": " << (sim->source_probe_oblique ? "\"fiber\"" : "\"ring\"") << ",\n";
        f << "\"num_SDS\": " << S << ",\n";
        f << "\"SDS_detected_number\": [";
        for (int s = 0; s < S - 1; s++) f << " " << (long long)res->detected[g0 + s] << " ,";
        f << " " << (long long)res->detected[g0 + S - 1] << " ]\n";
        f << "}";
    }
    return MCML_OK;
}

extern "C" int mcml_write_outputs(const mcml_sim* sim, const mcml_opts* opts, const mcml_result* res, int set,
                                  const char* output_txt, const char* outdir, const mcml_summary_info* info)
{
    if (!sim || !opts || !res || !output_txt || set < 0 || set >= res->nsets) return MCML_ERR_ARG;
    return write_outputs(sim, opts, res, set, output_txt, outdir, info, true);
}

/* ====================================================================== CLI */
namespace {

void show_usage(const std::string& name)
{
    size_t found = name.find_last_of("/\\");
    std::string exe = found == std::string::npos ? name : name.substr(found + 1);
    std::cerr << "Usage: " << exe << " sim_set.json input.txt output.txt <option(s)>\n"
              << "Options:\n"
              << "\t-h,--help\t Show this help message\n"
              << "\t-R,\t\t Replay the detected photon after first simulation\n"
              << "\t-A,\t\t Output the absorbance array\n"
              << "\t-P,\t\t Output the pathlength for each photon, otherwise output the calculated average pathlength\n"
              << "\t-AP,\t\t Calaulate and output the average pathlength\n"
              << "\t-B,\t\t Output the pathlength and absorbance file in binary format\n"
              << "\t-LS,\t\t List the GPUs and not run simulation\n"
              << "\t-G,\t\t Select which GPU to run, following by a number start from 0\n"
              << "\t-S,\t\t Random seed (default: time(NULL))\n"
              << std::endl;
}

int list_gpus(bool print)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        std::cout << "ERROR: No HIP-capable GPU device found\n";
        return 0;
    }
    for (int d = 0; d < n && print; d++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
        std::cout << "=======================   GPU Infomation  ==========================\n";
        printf("Device %d of %d:\t\t%s\n", d + 1, n, p.name);
        printf("Architecture:\t\t%s\n", p.gcnArchName);
        printf("Global Memory:\t\t%zu B\nShared Memory:\t\t%zu B\nClock Speed:\t\t%.2f GHz\n", (size_t)p.totalGlobalMem,
               (size_t)p.sharedMemPerBlock, p.clockRate * 1e-6);
        printf("Number of CUs:\t\t%d\nWavefront size:\t\t%d\n\n", p.multiProcessorCount, p.warpSize);
    }
    return n;
}

double steady_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* The executable's -P files, written on a thread of their own while the batch is read back:
 * group (set, SDS) is written once its rows have landed, in set-then-SDS order, which is the
 * order of the rows and the order in which the text files are appended to. */
struct PathlengthWriter {
    std::mutex m;
    std::condition_variable cv;
    uint64_t rows = 0;      /* rows of the result on the host */
    bool have_rows = false; /* the result's pl_offset is complete */
    bool finished = false;  /* the batch call has returned */
    int run_rc = MCML_OK;   /* ... with this code */
    bool done = false;      /* the writer thread reads the result no more */
    int rc = MCML_OK;       /* the writer's own result */
};

static void pathlength_rows_landed(void* user, const mcml_result*, uint64_t rows)
{
    PathlengthWriter* w = static_cast<PathlengthWriter*>(user);
    {
        std::lock_guard<std::mutex> lk(w->m);
        w->rows = rows;
        w->have_rows = true;
    }
    w->cv.notify_all();
}

/* The batch is about to free the result after an error: stop the writer and wait for it. */
static void pathlength_release(void* user)
{
    PathlengthWriter* w = static_cast<PathlengthWriter*>(user);
    std::unique_lock<std::mutex> lk(w->m);
    w->finished = true;
    w->run_rc = MCML_ERR_DEVICE;
    w->cv.notify_all();
    w->cv.wait(lk, [&] { return w->done; });
}

static void write_pathlength_groups(PathlengthWriter* w, const mcml_opts* o, const mcml_result* res, int nsets, int S)
{
    for (int j = 0; j < nsets; j++)
        for (int s = 0; s < S; s++) {
            {
                std::unique_lock<std::mutex> lk(w->m);
                const size_t g = (size_t)j * (size_t)S + (size_t)s;
                w->cv.wait(lk, [&] { return w->finished || (w->have_rows && w->rows >= res->pl_offset[g + 1]); });
                if (w->finished && w->run_rc != MCML_OK) return;
            }
            const int rc = write_pathlength_file(o, res, j, s, nullptr);
            if (rc != MCML_OK) {
                w->rc = rc;
                return;
            }
        }
}

static void write_pathlength_stream(PathlengthWriter* w, const mcml_opts* o, const mcml_result* res, int nsets,
                                    int S)
{
    write_pathlength_groups(w, o, res, nsets, S);
    {
        std::lock_guard<std::mutex> lk(w->m);
        w->done = true;
    }
    w->cv.notify_all();
}

/* Consecutive runs of input lines (first, count) whose absorbance grids fit the budget; a
 * line that alone exceeds it runs by itself. */
static std::vector<std::pair<int, int>> plan_batches(const mcml_sim* sims, int ns, const mcml_opts& o)
{
    double budget = 0.0;
    if (const char* e = getenv("MCML_BATCH_BUDGET_BYTES")) budget = strtod(e, nullptr);
    if (budget <= 0.0) {
        size_t free_b = 0, total_b = 0;
        budget = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? 0.5 * (double)free_b : 8e9;
    }
    std::vector<std::pair<int, int>> out;
    int b0 = 0;
    double used = 0.0;
    for (int i = 0; i < ns; i++) {
        const mcml_sim& m = sims[i];
        const double grid = o.replay && o.absorbance
                                ? 8.0 * (double)m.num_sds * (double)m.grid_nz * ((double)m.grid_nr + 1.0)
                                : 0.0;
        if (i > b0 && used + grid > budget) {
            out.push_back({b0, i - b0});
            b0 = i;
            used = 0.0;
        }
        used += grid;
    }
    if (ns > b0) out.push_back({b0, ns - b0});
    return out;
}

extern "C" int mcml_main(int argc, char** argv)
{
    mcml_opts o;
    memset(&o, 0, sizeof(o));
    int gpu = 0;
    uint64_t seed = (uint64_t)time(nullptr);
    bool seed_given = false;
    for (int i = 1; i < argc; i++) { /* main.cu:31-66 */
        std::string a(argv[i]);
        if (a == "-h") {
            show_usage(argv[0]);
            return 0;
        } else if (a == "-R") o.replay = 1;
        else if (a == "-A") o.absorbance = 1;
        else if (a == "-P") o.pathlength = 1;
        else if (a == "-AP") o.avg_pathlength = 1;
        else if (a == "-B") o.binary = 1;
        else if (a == "-G") {
            if (i + 1 >= argc) {
                std::cout << "GPU selection ERROR!\n";
                return 0;
            }
            int sel = atoi(argv[i + 1]);
            if (sel < 0) {
                std::cout << "GPU selection ERROR!\n";
                return 0;
            }
            gpu = sel;
        } else if (a == "-S") {
            if (i + 1 >= argc) {
                std::cout << "-S needs a seed\n";
                return 0;
            }
            seed = strtoull(argv[i + 1], nullptr, 10);
            seed_given = true;
        } else if (a == "-LS") {
            list_gpus(true);
            return 0;
        }
    }
    if (argc < 4) {
        show_usage(argv[0]);
        return 0;
    }
    if (o.replay && !o.pathlength) o.avg_pathlength = 1; /* main.cu:73-75 */
    if (o.pathlength && !o.replay) {
        std::cout << "-P option only work with -R option!\n";
        return 0;
    }
    if (o.absorbance && !o.replay) {
        std::cout << "-A option only work with -R option!\n";
        return 0;
    }
    if (o.avg_pathlength && !o.replay) {
        std::cout << "-AP option only work with -R option!\n";
        return 0;
    }
    if (o.binary && !(o.pathlength || o.absorbance)) {
        std::cout << "-B option only work with -P or -A option!\n";
        return 0;
    }
    o.skip_records = 1; /* the files need the sums, rows and grids, not the per-photon records */
    mcml_sim* sims = nullptr;
    int ns = 0;
    char msg[1024];
    int rc = mcml_read_input(argv[1], argv[2], &sims, &ns, msg, sizeof(msg));
    if (rc != MCML_OK) {
        std::cout << msg;
        printf("Something wrong with read_simulation_data!\n");
        return 1;
    }
    std::cout << "******************************" << std::endl;
    std::cout << "*       MCML_GPU K" << kVersion << "       *" << std::endl;
    std::cout << "*         " << kUpdated << "         *" << std::endl;
    std::cout << "******************************" << std::endl;
    int ngpu = list_gpus(true);
    if (gpu >= ngpu) {
        std::cout << "GPU selection ERROR!\n";
        mcml_free_sims(sims);
        return 0;
    }
    std::cout << "GPU " << gpu << " is selected\n";
    if (!seed_given) printf("seed %llu\n", (unsigned long long)seed);
    int err = 0;
    mcml_ctx* ctx = mcml_create(gpu, &err);
    if (!ctx) {
        printf("Error code=%i, cannot open GPU %d\n", err, gpu);
        mcml_free_sims(sims);
        return 1;
    }
    const double t0 = steady_s();
    /* The lines of input.txt run as few persistent launches as memory allows: consecutive
     * lines are batched while their absorbance grids (held on the device and again in the
     * result) fit half of the device memory that is free, or MCML_BATCH_BUDGET_BYTES when that
     * is set.  Photon streams depend on (seed, set, photon) through seed + 2^64/phi (set + 1)
     * only, so running lines b.. as sets 0.. under seed + 2^64/phi b gives every line the
     * stream it has in a single batch: the outputs do not depend on the split.  Outputs are
     * written per simulation in input order, each batch before the next runs. */
    const std::vector<std::pair<int, int>> batches = plan_batches(sims, ns, o);
    for (size_t bi = 0; bi < batches.size() && rc == MCML_OK; bi++) {
        const int b0 = batches[bi].first, nb = batches[bi].second;
        mcml_result res;
        /* -P: the pathlength files are written while the rest is read back */
        const bool stream_pl = o.replay && o.pathlength;
        PathlengthWriter pw;
        mcml_readback_sink sink = {pathlength_rows_landed, pathlength_release, &pw};
        std::thread writer;
        bool pl_written = false; /* the writer thread has written (or will have written) the -P files */
        if (stream_pl) {
            try {
                writer = std::thread(write_pathlength_stream, &pw, &o, &res, nb, sims[b0].num_sds);
                pl_written = true;
            } catch (...) { /* no thread to spare: the files are written after the batch */
            }
        }
        rc = mcml_run_batch_sink(ctx, sims + b0, nb, &o, seed + 0x9E3779B97F4A7C15ull * (uint64_t)b0, &res,
                                 pl_written ? &sink : nullptr);
        if (pl_written) {
            {
                std::lock_guard<std::mutex> lk(pw.m);
                pw.finished = true;
                pw.run_rc = rc;
            }
            pw.cv.notify_all();
            writer.join();
        }
        if (rc != MCML_OK) {
            printf("Error code=%i, %s.\n", rc, mcml_last_error(ctx));
            break;
        }
        if (pw.rc != MCML_OK) {
            rc = pw.rc;
            printf("Error code=%i writing outputs\n", rc);
            mcml_result_free(&res);
            break;
        }
        uint64_t total = 0;
        for (int i = b0; i < b0 + nb; i++) total += sims[i].number_photons;
        for (int i = b0; i < b0 + nb; i++) {
            const int j = i - b0;
            printf("simulating %d\n", i);
            printf("photon deteced number for SDSs:");
            for (int s = 0; s < res.num_sds; s++)
                printf("\t%llu,", (unsigned long long)res.detected[(size_t)j * res.num_sds + s]);
            printf("\n");
            const double share = total ? (double)sims[i].number_photons / (double)total : 0.0;
            mcml_summary_info info;
            info.sim_time_s = res.ms_first_pass_wall * 1e-3 * share;
            info.replay_time_s = res.ms_replay_wall * 1e-3 * share;
            info.gpu_name = mcml_device_name(ctx);
            printf("finish first run, cost %.2f secs\n", info.sim_time_s);
            rc = write_outputs(&sims[i], &o, &res, j, argv[3], nullptr, &info, !pl_written);
            if (rc != MCML_OK) {
                printf("Error code=%i writing outputs\n", rc);
                break;
            }
            if (o.replay) printf("finish replay, cost %.2f secs\n", info.replay_time_s);
        }
        mcml_result_free(&res);
    }
    printf("Simulation time: %.2f sec\n", steady_s() - t0);
    mcml_destroy(ctx);
    mcml_free_sims(sims);
    return rc == MCML_OK ? 0 : 1;
}
