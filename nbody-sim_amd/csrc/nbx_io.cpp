// nbx_io.cpp -- host-side run support behind the C ABI: the JSON config of
// Config.h / Config.cpp, the disk generator of DiskInit.cpp, and the
// frames.bin (version 2) / stats.csv writer of OutputWriter.cpp, plus the
// visualizer's perturber.json sidecar.  Plain host C++ (no device code).
//
// File formats (byte-for-byte those of OutputWriter.cpp:44-197):
//   frames.bin header, 32 bytes, little endian:
//     "NBOD" | uint32 version = 2 | uint32 n_particles | uint32 n_frames
//     (patched by finalize) | double dt | double dt * output_every_n_steps
//   frame: double time_T0 | float x[n] | float y[n] | float z[n] |
//          uint16 agg_size[n]            -> 8 + 14 n bytes per frame
//   stats.csv: the header line below, then one row per aggregate with the
//   numbers printed as an ostream prints them (floats: %g, 6 significant
//   digits), and the perturber row with agg_id -1.
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <cctype>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/nbx.h"

namespace nbx {
void set_last_error(const char *msg);  // nbx_abi.hip: the message of nbx_last_error()
}

namespace {

int io_fail(int code, const char *fmt, ...)
{
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    nbx::set_last_error(buf);
    return code;
}

// ---------------------------------------------------------------------------
// A small JSON reader: one object whose values are numbers, strings, booleans
// or null; nested objects / arrays are parsed and skipped.
// ---------------------------------------------------------------------------
struct JVal {
    enum Kind { kNum, kStr, kBool, kNull, kOther } kind = kNull;
    double num = 0;
    bool b = false;
    std::string str;
};

struct JParser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit JParser(const std::string &src) : s(src) {}

    void ws()
    {
        while (i < s.size() && isspace((unsigned char)s[i])) ++i;
    }
    bool fail(const char *what)
    {
        if (err.empty()) err = std::string(what) + " at byte " + std::to_string(i);
        return false;
    }
    bool lit(const char *w)
    {
        const size_t n = strlen(w);
        if (s.compare(i, n, w) != 0) return fail("unexpected token");
        i += n;
        return true;
    }
    bool string(std::string &out)
    {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        ++i;
        out.clear();
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) return fail("bad escape");
                const char e = s[i++];
                switch (e) {
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case 'r': c = '\r'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'u': {  // keep ASCII code points, replace the rest
                    if (i + 4 > s.size()) return fail("bad \\u escape");
                    const long cp = strtol(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
                    c = cp < 128 ? (char)cp : '?';
                    break;
                }
                default: c = e; break;
                }
            }
            out.push_back(c);
        }
        if (i >= s.size()) return fail("unterminated string");
        ++i;
        return true;
    }
    bool number(double &out)
    {
        const char *b = s.c_str() + i;
        char *e = nullptr;
        out = strtod(b, &e);
        if (e == b) return fail("expected number");
        i += (size_t)(e - b);
        return true;
    }
    bool value(JVal &v)
    {
        ws();
        if (i >= s.size()) return fail("unexpected end");
        const char c = s[i];
        if (c == '"') {
            v.kind = JVal::kStr;
            return string(v.str);
        }
        if (c == 't') {
            v.kind = JVal::kBool;
            v.b = true;
            return lit("true");
        }
        if (c == 'f') {
            v.kind = JVal::kBool;
            v.b = false;
            return lit("false");
        }
        if (c == 'n') {
            v.kind = JVal::kNull;
            return lit("null");
        }
        if (c == '{' || c == '[') {
            v.kind = JVal::kOther;
            return skip_compound();
        }
        v.kind = JVal::kNum;
        return number(v.num);
    }
    bool skip_compound()
    {
        const char open = s[i], close = open == '{' ? '}' : ']';
        ++i;
        ws();
        if (i < s.size() && s[i] == close) {
            ++i;
            return true;
        }
        while (true) {
            if (open == '{') {
                std::string k;
                ws();
                if (!string(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
            }
            JVal v;
            if (!value(v)) return false;
            ws();
            if (i < s.size() && s[i] == ',') {
                ++i;
                continue;
            }
            if (i < s.size() && s[i] == close) {
                ++i;
                return true;
            }
            return fail("expected ',' or closing bracket");
        }
    }
    bool object(std::vector<std::pair<std::string, JVal>> &out)
    {
        ws();
        if (i >= s.size() || s[i] != '{') return fail("expected '{'");
        ++i;
        ws();
        if (i < s.size() && s[i] == '}') {
            ++i;
        } else {
            while (true) {
                std::string k;
                ws();
                if (!string(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
                JVal v;
                if (!value(v)) return false;
                out.emplace_back(k, v);
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    break;
                }
                return fail("expected ',' or '}'");
            }
        }
        ws();
        if (i != s.size()) return fail("trailing characters");
        return true;
    }
};

// ---------------------------------------------------------------------------
// Counter-based generator: draw k of body i under `seed` (splitmix64
// finaliser, Steele, Lea & Flood 2014), uniform in [0, 1) with 24 bits.
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct BodyRng {
    uint64_t key;
    uint64_t k = 0;
    BodyRng(uint32_t seed, int32_t i) : key(mix64(((uint64_t)seed << 32) | (uint32_t)i)) {}
    float uniform() { return (float)(mix64(key + k++) >> 40) * (1.0f / 16777216.0f); }
    float normal()  // Box-Muller
    {
        const float u1 = 1.0f - uniform(), u2 = uniform();
        return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717959f * u2);
    }
    float rayleigh(float sigma) { return sigma * sqrtf(-2.0f * logf(1.0f - uniform())); }
};

std::string fmt_g(float v)
{
    char b[64];
    snprintf(b, sizeof(b), "%g", (double)v);
    return b;
}

}  // namespace

struct nbx_writer {
    FILE *frames = nullptr;
    FILE *stats = nullptr;
    int32_t n = 0;
};

namespace {

// OutputWriter.h:10-12
const float kT0ToYears = 1.0f / (2.0f * 3.14159265f);
const float kMsunToMearth = 332946.0f;

int put(FILE *f, const void *p, size_t bytes)
{
    if (fwrite(p, 1, bytes, f) != bytes) return io_fail(NBX_E_ARG, "write failed: %s", strerror(errno));
    return NBX_OK;
}

int put_row(nbx_writer *w, const std::string &row)
{
    if (fputs(row.c_str(), w->stats) < 0) return io_fail(NBX_E_ARG, "stats.csv write failed");
    return NBX_OK;
}

}  // namespace

extern "C" {

int nbx_config_default(nbx_config *c)
{
    if (!c) return io_fail(NBX_E_ARG, "null config");
    memset(c, 0, sizeof(*c));
    c->n_particles = 100000;
    c->disk_r_min = 0.5f;
    c->disk_r_max = 5.0f;
    c->disk_h_factor = 0.05f;
    c->particle_mass_msun = 3.003e-7f;
    c->star_mass_msun = 1.0f;
    c->softening_AU = 0.01f;
    c->theta = 0.5f;
    c->collision_radius_factor = 100.0f;
    c->restitution = 0.3f;
    c->gas_drag_rate = 0.0f;
    c->dt = 0.01f;
    c->total_time = 2000.0f;
    c->output_every_n_steps = 200;
    snprintf(c->output_dir, sizeof(c->output_dir), "%s", "data");
    c->perturber_enabled = 0;
    c->perturber_mass_mjup = 1.0f;
    c->perturber_radius_AU = 20.0f;
    c->perturber_phase0 = 0.0f;
    c->disk_seed = 42u;
    c->e_sigma = -1.0f;
    c->i_sigma = -1.0f;
    return NBX_OK;
}

int nbx_config_load(const char *path, nbx_config *c)
{
    if (!path || !c) return io_fail(NBX_E_ARG, "null argument");
    nbx_config_default(c);
    FILE *f = fopen(path, "rb");
    if (!f) return io_fail(NBX_E_ARG, "Cannot open config: %s", path);
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    fclose(f);
    std::vector<std::pair<std::string, JVal>> kv;
    JParser P(text);
    if (!P.object(kv))
        return io_fail(NBX_E_ARG, "JSON parse error in '%s': %s", path, P.err.c_str());
    for (const auto &e : kv) {
        const std::string &k = e.first;
        const JVal &v = e.second;
        auto num = [&](float *dst) -> int {
            if (v.kind != JVal::kNum)
                return io_fail(NBX_E_ARG, "config '%s': %s must be a number", path, k.c_str());
            *dst = (float)v.num;
            return NBX_OK;
        };
        auto inum = [&](int32_t *dst) -> int {
            if (v.kind != JVal::kNum)
                return io_fail(NBX_E_ARG, "config '%s': %s must be a number", path, k.c_str());
            *dst = (int32_t)v.num;
            return NBX_OK;
        };
        int rc = NBX_OK;
        if (k == "n_particles") rc = inum(&c->n_particles);
        else if (k == "disk_r_min") rc = num(&c->disk_r_min);
        else if (k == "disk_r_max") rc = num(&c->disk_r_max);
        else if (k == "disk_h_factor") rc = num(&c->disk_h_factor);
        else if (k == "particle_mass_msun") rc = num(&c->particle_mass_msun);
        else if (k == "star_mass_msun") rc = num(&c->star_mass_msun);
        else if (k == "softening_AU") rc = num(&c->softening_AU);
        else if (k == "theta") rc = num(&c->theta);
        else if (k == "collision_radius_factor") rc = num(&c->collision_radius_factor);
        else if (k == "restitution") rc = num(&c->restitution);
        else if (k == "gas_drag_rate") rc = num(&c->gas_drag_rate);
        else if (k == "dt") rc = num(&c->dt);
        else if (k == "total_time") rc = num(&c->total_time);
        else if (k == "output_every_n_steps") rc = inum(&c->output_every_n_steps);
        else if (k == "perturber_mass_mjup") rc = num(&c->perturber_mass_mjup);
        else if (k == "perturber_radius_AU") rc = num(&c->perturber_radius_AU);
        else if (k == "perturber_phase0") rc = num(&c->perturber_phase0);
        else if (k == "e_sigma") rc = num(&c->e_sigma);
        else if (k == "i_sigma") rc = num(&c->i_sigma);
        else if (k == "disk_seed") {
            if (v.kind != JVal::kNum)
                return io_fail(NBX_E_ARG, "config '%s': disk_seed must be a number", path);
            c->disk_seed = (uint32_t)v.num;
        } else if (k == "output_dir") {
            if (v.kind != JVal::kStr)
                return io_fail(NBX_E_ARG, "config '%s': output_dir must be a string", path);
            if (v.str.size() >= sizeof(c->output_dir))
                return io_fail(NBX_E_ARG, "config '%s': output_dir too long", path);
            snprintf(c->output_dir, sizeof(c->output_dir), "%s", v.str.c_str());
        } else if (k == "perturber_enabled") {
            if (v.kind != JVal::kBool)
                return io_fail(NBX_E_ARG, "config '%s': perturber_enabled must be a boolean",
                               path);
            c->perturber_enabled = v.b ? 1 : 0;
        }
        if (rc) return rc;
    }
    // Config.cpp:40-52, in the same order
    if (c->perturber_enabled) {
        if (c->perturber_mass_mjup <= 0.0f)
            return io_fail(NBX_E_ARG, "perturber_mass_mjup must be positive");
        if (c->perturber_radius_AU <= 0.0f)
            return io_fail(NBX_E_ARG, "perturber_radius_AU must be positive");
    }
    if (c->n_particles <= 0)
        return io_fail(NBX_E_ARG, "n_particles must be positive, got: %d", c->n_particles);
    if (c->dt <= 0.f) return io_fail(NBX_E_ARG, "dt must be positive");
    if (c->disk_r_min >= c->disk_r_max)
        return io_fail(NBX_E_ARG, "disk_r_min must be less than disk_r_max");
    return NBX_OK;
}

// DiskInit.cpp:10-16 in float.
float nbx_particle_radius(const nbx_config *c)
{
    if (!c) return 0.f;
    const float pi = 3.14159265358979f, au = 1.496e11f, msun = 1.989e30f;
    const float mass_kg = c->particle_mass_msun * msun;
    const float vol = mass_kg / 3000.0f;
    const float r_m = cbrtf(3.0f * vol / (4.0f * pi));
    return r_m / au;
}

int nbx_init_disk(const nbx_config *c, uint32_t seed, nbx_particles *p)
{
    if (!c || !p) return io_fail(NBX_E_ARG, "null argument");
    if (c->n_particles <= 0) return io_fail(NBX_E_ARG, "initDisk: n_particles must be positive");
    if (c->disk_r_min <= 0.0f)
        return io_fail(NBX_E_ARG, "initDisk: disk_r_min must be positive");
    if (c->disk_r_min >= c->disk_r_max)
        return io_fail(NBX_E_ARG, "initDisk: disk_r_min must be less than disk_r_max");
    if (p->n != c->n_particles)
        return io_fail(NBX_E_ARG, "initDisk: arrays hold %d bodies, config %d", p->n,
                       c->n_particles);
    if (!p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->ax || !p->ay || !p->az ||
        !p->mass)
        return io_fail(NBX_E_ARG, "initDisk: null particle array");
    const float pi = 3.14159265358979f;
    const float srmin = sqrtf(c->disk_r_min), srmax = sqrtf(c->disk_r_max);
    const bool rayleigh = c->e_sigma >= 0.0f || c->i_sigma >= 0.0f;
    for (int32_t i = 0; i < c->n_particles; ++i) {
        BodyRng g(seed, i);
        // Sigma ~ r^-3/2: p(r) ~ r^-1/2, inverse CDF in sqrt(r) (DiskInit.cpp:39-44)
        const float sr = g.uniform() * (srmax - srmin) + srmin;
        const float r = sr * sr;
        const float phi = g.uniform() * 2.0f * pi;
        const float cph = cosf(phi), sph = sinf(phi);
        p->x[i] = r * cph;
        p->z[i] = r * sph;
        p->y[i] = g.normal() * (c->disk_h_factor * r);
        const float vc = sqrtf(c->star_mass_msun / r);
        if (!rayleigh) {  // DiskInit.cpp:55-60: 1 % Gaussian noise on a circular orbit
            const float ns = 0.01f * vc;
            p->vx[i] = -vc * sph + g.normal() * ns;
            p->vz[i] = vc * cph + g.normal() * ns;
            p->vy[i] = g.normal() * ns;
        } else {  // epicyclic e and vertical i of Rayleigh width
            const float e = g.rayleigh(c->e_sigma > 0.f ? c->e_sigma : 0.f);
            const float psi = g.uniform() * 2.0f * pi;
            const float inc = g.rayleigh(c->i_sigma > 0.f ? c->i_sigma : 0.f);
            const float om = g.uniform() * 2.0f * pi;
            const float v_r = e * vc * sinf(psi);
            const float v_t = vc * (1.0f + 0.5f * e * cosf(psi));
            p->vx[i] = v_r * cph - v_t * sph;
            p->vz[i] = v_r * sph + v_t * cph;
            p->vy[i] = inc * vc * cosf(om);
        }
        p->ax[i] = p->ay[i] = p->az[i] = 0.0f;
        p->mass[i] = c->particle_mass_msun;
        if (p->id) p->id[i] = i;
    }
    return NBX_OK;
}

int nbx_writer_open(nbx_writer **out, const char *dir, int32_t n, float dt, int32_t every)
{
    if (!out || !dir) return io_fail(NBX_E_ARG, "null argument");
    *out = nullptr;
    if (n <= 0) return io_fail(NBX_E_ARG, "writer: n_particles must be positive");
    nbx_writer *w = new nbx_writer();
    w->n = n;
    const std::string fp = std::string(dir) + "/frames.bin", sp = std::string(dir) + "/stats.csv";
    w->frames = fopen(fp.c_str(), "wb");
    if (!w->frames) {
        delete w;
        return io_fail(NBX_E_ARG, "OutputWriter: cannot open %s", fp.c_str());
    }
    w->stats = fopen(sp.c_str(), "w");
    if (!w->stats) {
        fclose(w->frames);
        delete w;
        return io_fail(NBX_E_ARG, "OutputWriter: cannot open %s", sp.c_str());
    }
    unsigned char hdr[32];
    const uint32_t version = 2u, np = (uint32_t)n, nf = 0u;
    const double dtd = (double)dt, odt = (double)dt * (double)every;
    memcpy(hdr, "NBOD", 4);
    memcpy(hdr + 4, &version, 4);
    memcpy(hdr + 8, &np, 4);
    memcpy(hdr + 12, &nf, 4);
    memcpy(hdr + 16, &dtd, 8);
    memcpy(hdr + 24, &odt, 8);
    int rc = put(w->frames, hdr, sizeof(hdr));
    if (!rc)
        rc = put_row(w, "frame,sim_time_T0,sim_time_years,n_aggregates,agg_id,n_particles,"
                        "mass_Msun,mass_Mearth,cx_AU,cy_AU,cz_AU,vx,vy,vz,dist_center_AU\n");
    if (rc) {
        nbx_writer_close(w);
        return rc;
    }
    *out = w;
    return NBX_OK;
}

int nbx_writer_frame(nbx_writer *w, const nbx_particles *p, float t, const uint16_t *agg)
{
    if (!w || !p || !agg) return io_fail(NBX_E_ARG, "null argument");
    if (p->n != w->n)
        return io_fail(NBX_E_ARG, "OutputWriter::writeFrame: agg_sizes_per_particle size mismatch");
    const double td = (double)t;
    int rc = put(w->frames, &td, 8);
    if (!rc) rc = put(w->frames, p->x, sizeof(float) * w->n);
    if (!rc) rc = put(w->frames, p->y, sizeof(float) * w->n);
    if (!rc) rc = put(w->frames, p->z, sizeof(float) * w->n);
    if (!rc) rc = put(w->frames, agg, sizeof(uint16_t) * w->n);
    return rc;
}

int nbx_writer_stats(nbx_writer *w, int32_t frame, float t, int32_t n_agg,
                     const nbx_aggregate *a)
{
    if (!w || (n_agg > 0 && !a)) return io_fail(NBX_E_ARG, "null argument");
    const float years = t * kT0ToYears;
    for (int k = 0; k < n_agg; ++k) {
        const nbx_aggregate &g = a[k];
        std::string row = std::to_string(frame) + ',' + fmt_g(t) + ',' + fmt_g(years) + ',' +
                          std::to_string(n_agg) + ',' + std::to_string(g.id) + ',' +
                          std::to_string(g.n_particles) + ',' + fmt_g(g.mass_msun) + ',' +
                          fmt_g(g.mass_msun * kMsunToMearth) + ',' + fmt_g(g.cx) + ',' +
                          fmt_g(g.cy) + ',' + fmt_g(g.cz) + ',' + fmt_g(g.vx) + ',' +
                          fmt_g(g.vy) + ',' + fmt_g(g.vz) + ',' + fmt_g(g.dist_center) + '\n';
        int rc = put_row(w, row);
        if (rc) return rc;
    }
    return NBX_OK;
}

int nbx_writer_perturber(nbx_writer *w, int32_t frame, float t, int32_t n_agg, float px,
                         float py, float pz, float mass_msun, float omega)
{
    if (!w) return io_fail(NBX_E_ARG, "null writer");
    const float years = t * kT0ToYears;
    const float r = sqrtf(px * px + pz * pz);
    // tangent in the XZ plane, (-z, 0, x) / r, speed omega r
    const float v_mag = omega * r;
    const float vx = (r > 0.f) ? (-pz / r) * v_mag : 0.f;
    const float vz = (r > 0.f) ? (px / r) * v_mag : 0.f;
    const float vy = 0.f;
    std::string row = std::to_string(frame) + ',' + fmt_g(t) + ',' + fmt_g(years) + ',' +
                      std::to_string(n_agg) + ",-1,0," + fmt_g(mass_msun) + ',' +
                      fmt_g(mass_msun * kMsunToMearth) + ',' + fmt_g(px) + ',' + fmt_g(py) +
                      ',' + fmt_g(pz) + ',' + fmt_g(vx) + ',' + fmt_g(vy) + ',' + fmt_g(vz) +
                      ',' + fmt_g(r) + '\n';
    return put_row(w, row);
}

int nbx_writer_finalize(nbx_writer *w, int32_t frames_written)
{
    if (!w) return io_fail(NBX_E_ARG, "null writer");
    const uint32_t nf = (uint32_t)frames_written;
    const long end = ftell(w->frames);
    if (fseek(w->frames, 12, SEEK_SET) != 0) return io_fail(NBX_E_ARG, "frames.bin seek failed");
    int rc = put(w->frames, &nf, 4);
    if (fseek(w->frames, end, SEEK_SET) != 0 && !rc) rc = io_fail(NBX_E_ARG, "seek failed");
    fflush(w->frames);
    fflush(w->stats);
    return rc;
}

int nbx_writer_close(nbx_writer *w)
{
    if (!w) return NBX_OK;
    if (w->frames) fclose(w->frames);
    if (w->stats) fclose(w->stats);
    delete w;
    return NBX_OK;
}

int nbx_write_perturber_sidecar(const char *dir, int32_t enabled, float mass_msun, float r,
                                float phase0, float omega)
{
    if (!dir) return io_fail(NBX_E_ARG, "null directory");
    const std::string path = std::string(dir) + "/perturber.json";
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return io_fail(NBX_E_ARG, "cannot open %s", path.c_str());
    fprintf(f,
            "{\n  \"enabled\": %s,\n  \"mass_msun\": %.9g,\n  \"r_AU\": %.9g,\n"
            "  \"phase0_rad\": %.9g,\n  \"omega_T0_inv\": %.9g\n}\n",
            enabled ? "true" : "false", (double)mass_msun, (double)r, (double)phase0,
            (double)omega);
    fclose(f);
    return NBX_OK;
}

}  // extern "C"
