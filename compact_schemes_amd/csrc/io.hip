// io.hip -- host side of a run: the YAML run configuration, the block grid
// and the legacy-VTK writer (no device code).
//
// The reference reads config.yaml with yaml-cpp (driver/canard.cu:37-98;
// domdcomp.hpp:67-107, grid.hpp:74-87, physics_base.hpp:135-145), generates
// the block's point coordinates (grid.hpp:89-110) and writes big-endian
// float legacy VTK structured grids (IO.hpp:108-159).  yaml-cpp is not a
// dependency here: the schema is a mapping of sections, each a sequence of
// one-key mappings whose values are numbers or flow sequences of numbers, and
// this file reads exactly that.
#include "cs_internal.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cs {
namespace {

struct YamlEntry {
    std::string key;
    std::vector<std::string> values;   // one for a scalar, several for [a, b, c]
    bool list = false;
    int line = 0;
};
typedef std::map<std::string, std::vector<YamlEntry>> YamlDoc;

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

int yaml_error(const std::string& path, int line, const std::string& what) {
    return fail(CS_ERR_ARG, path + ":" + std::to_string(line) + ": " + what);
}

// sections at column 0 ("name:"), entries "- key: value" or "key: value"
// indented below them; '#' starts a comment
int parse_yaml(const std::string& path, YamlDoc& doc) {
    std::ifstream in(path);
    if (!in) return fail(CS_ERR_ARG, "cannot open run configuration " + path);
    std::string raw, section;
    int line = 0;
    while (std::getline(in, raw)) {
        line++;
        size_t hash = raw.find('#');
        if (hash != std::string::npos) raw = raw.substr(0, hash);
        if (trim(raw).empty()) continue;
        const bool indented = raw[0] == ' ' || raw[0] == '\t';
        std::string body = trim(raw);
        if (!indented) {
            if (body.back() != ':' || body.find(':') != body.size() - 1)
                return yaml_error(path, line, "expected a section name followed by ':'");
            section = trim(body.substr(0, body.size() - 1));
            if (section.empty()) return yaml_error(path, line, "empty section name");
            doc[section];
            continue;
        }
        if (section.empty()) return yaml_error(path, line, "entry outside a section");
        if (body[0] == '-') body = trim(body.substr(1));
        size_t colon = body.find(':');
        if (colon == std::string::npos || colon == 0) return yaml_error(path, line, "expected 'key: value'");
        YamlEntry e;
        e.key = trim(body.substr(0, colon));
        e.line = line;
        std::string val = trim(body.substr(colon + 1));
        if (val.empty()) return yaml_error(path, line, "missing value of " + e.key);
        if (val[0] == '[') {
            if (val.back() != ']') return yaml_error(path, line, "unterminated sequence of " + e.key);
            e.list = true;
            std::stringstream ss(val.substr(1, val.size() - 2));
            std::string item;
            while (std::getline(ss, item, ',')) {
                item = trim(item);
                if (item.empty()) return yaml_error(path, line, "empty item in the sequence of " + e.key);
                e.values.push_back(item);
            }
        } else {
            e.values.push_back(val);
        }
        doc[section].push_back(e);
    }
    return CS_OK;
}

const YamlEntry* find_entry(const YamlDoc& doc, const char* section, const char* key) {
    auto it = doc.find(section);
    if (it == doc.end()) return nullptr;
    for (const YamlEntry& e : it->second)
        if (e.key == key) return &e;
    return nullptr;
}

int to_double(const std::string& path, const YamlEntry& e, const std::string& s, double* out) {
    char* end = nullptr;
    errno = 0;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE || !std::isfinite(v))
        return yaml_error(path, e.line, "not a number: '" + s + "' (" + e.key + ")");
    *out = v;
    return CS_OK;
}

int to_int(const std::string& path, const YamlEntry& e, const std::string& s, int* out) {
    char* end = nullptr;
    errno = 0;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE || v < INT32_MIN || v > INT32_MAX)
        return yaml_error(path, e.line, "not an integer: '" + s + "' (" + e.key + ")");
    *out = (int)v;
    return CS_OK;
}

// required (or optional) scalar / sequence readers
int get_double(const std::string& path, const YamlDoc& doc, const char* sec, const char* key, bool required,
               double* out) {
    const YamlEntry* e = find_entry(doc, sec, key);
    if (!e) return required ? fail(CS_ERR_ARG, path + ": missing " + sec + "." + key) : CS_OK;
    if (e->list || e->values.size() != 1) return yaml_error(path, e->line, std::string(key) + " must be a number");
    return to_double(path, *e, e->values[0], out);
}

int get_int(const std::string& path, const YamlDoc& doc, const char* sec, const char* key, bool required, int* out) {
    const YamlEntry* e = find_entry(doc, sec, key);
    if (!e) return required ? fail(CS_ERR_ARG, path + ": missing " + sec + "." + key) : CS_OK;
    if (e->list || e->values.size() != 1) return yaml_error(path, e->line, std::string(key) + " must be an integer");
    return to_int(path, *e, e->values[0], out);
}

// a sequence of at least n integers; the first n are kept (block 0)
int get_ints(const std::string& path, const YamlDoc& doc, const char* sec, const char* key, int n, int* out) {
    const YamlEntry* e = find_entry(doc, sec, key);
    if (!e) return fail(CS_ERR_ARG, path + ": missing " + sec + "." + key);
    if ((int)e->values.size() < n)
        return yaml_error(path, e->line, std::string(key) + " needs " + std::to_string(n) + " values");
    for (int i = 0; i < n; i++) {
        int rc = to_int(path, *e, e->values[i], out + i);
        if (rc) return rc;
    }
    return CS_OK;
}

bool known_bc(int c) {
    return c == CS_BC_NON_REFLECTIVE || c == CS_BC_WALL_INVISCID || c == CS_BC_WALL_VISCOUS ||
           c == CS_BC_INTER_CURV || c == CS_BC_INTER_STRAIGHT || c == CS_BC_INTER_SUBDOMAINS || c == CS_BC_PERIODIC;
}

void put_be32(std::ofstream& out, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    unsigned char b[4] = {(unsigned char)(u >> 24), (unsigned char)(u >> 16), (unsigned char)(u >> 8),
                          (unsigned char)u};
    out.write(reinterpret_cast<const char*>(b), 4);
}

void put_be64(std::ofstream& out, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    unsigned char b[8];
    for (int i = 0; i < 8; i++) b[i] = (unsigned char)(u >> (56 - 8 * i));
    out.write(reinterpret_cast<const char*>(b), 8);
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" int cs_config_defaults(cs_run_config* cfg) {
    if (!cfg) return fail(CS_ERR_ARG, "null argument");
    std::memset(cfg, 0, sizeof(*cfg));
    for (int a = 0; a < 3; a++) cfg->nbpc[a] = 1;
    for (int i = 0; i < 6; i++) { cfg->nbbc[i] = CS_BC_NON_REFLECTIVE; cfg->mbcd[i] = -1; }
    // driver/canard.cu:117-127
    cfg->cfl = 0.95;
    cfg->tmax = 1.0;
    cfg->ndata = 2;
    cfg->nsteps = 0;
    cfg->viscous = 1;
    cfg->re = 1.0e4;
    cfg->alpha_f = 0.49;
    cfg->filter = 1;
    cfg->pulse_amp = 1.0e-3;
    cfg->pulse_width = 8.0;
    return CS_OK;
}

extern "C" int cs_config_load(const char* path, cs_run_config* cfg) {
    if (!path || !cfg) return fail(CS_ERR_ARG, "null argument");
    int rc = cs_config_defaults(cfg);
    if (rc) return rc;
    YamlDoc doc;
    const std::string p(path);
    if ((rc = parse_yaml(p, doc))) return rc;
    // driver (driver/canard.cu:38-39) and domdcomp (domdcomp.hpp:67-107): block 0
    if ((rc = get_int(p, doc, "driver", "nblocks", true, &cfg->nblocks))) return rc;
    if (cfg->nblocks < 0) return fail(CS_ERR_ARG, p + ": nblocks must be >= 0");
    const int nb = cfg->nblocks + 1;
    {
        // nbpc holds nbpc[axis][block] axis-major (domdcomp.hpp:69-76)
        const YamlEntry* e = find_entry(doc, "domdcomp", "nbpc");
        if (!e) return fail(CS_ERR_ARG, p + ": missing domdcomp.nbpc");
        if ((int)e->values.size() < 3 * nb) return yaml_error(p, e->line, "nbpc needs 3*(nblocks+1) values");
        for (int a = 0; a < 3; a++)
            if ((rc = to_int(p, *e, e->values[a * nb], &cfg->nbpc[a]))) return rc;
    }
    if ((rc = get_ints(p, doc, "domdcomp", "lximb", 1, &cfg->lximb))) return rc;
    if ((rc = get_ints(p, doc, "domdcomp", "letmb", 1, &cfg->letmb))) return rc;
    if ((rc = get_ints(p, doc, "domdcomp", "lzemb", 1, &cfg->lzemb))) return rc;
    {
        // nbbc / mbcd index mb + nn*(nblocks+1) + ip*3*(nblocks+1) (domdcomp.hpp:128-131)
        for (const char* key : {"nbbc", "mbcd"}) {
            const YamlEntry* e = find_entry(doc, "domdcomp", key);
            if (!e) return fail(CS_ERR_ARG, p + ": missing domdcomp." + key);
            if ((int)e->values.size() < 6 * nb) return yaml_error(p, e->line, std::string(key) + " needs 6*(nblocks+1) values");
            int* dst = key[0] == 'n' ? cfg->nbbc : cfg->mbcd;
            for (int ip = 0; ip < 2; ip++)
                for (int nn = 0; nn < 3; nn++)
                    if ((rc = to_int(p, *e, e->values[nn * nb + ip * 3 * nb], &dst[ip * 3 + nn]))) return rc;
        }
    }
    if (cfg->lximb < 1 || cfg->letmb < 1 || cfg->lzemb < 1)
        return fail(CS_ERR_SHAPE, p + ": lximb, letmb and lzemb must be >= 1 (points - 1)");
    for (int a = 0; a < 3; a++)
        if (cfg->nbpc[a] < 1) return fail(CS_ERR_ARG, p + ": nbpc values must be >= 1");
    for (int i = 0; i < 6; i++)
        if (!known_bc(cfg->nbbc[i]))
            return fail(CS_ERR_ARG, p + ": unknown boundary code " + std::to_string(cfg->nbbc[i]) + " in nbbc");
    // grid (grid.hpp:74-87) and physics (physics_base.hpp:135-145)
    if ((rc = get_double(p, doc, "grid", "doml0", true, &cfg->doml0))) return rc;
    if ((rc = get_double(p, doc, "grid", "doml1", true, &cfg->doml1))) return rc;
    if ((rc = get_double(p, doc, "grid", "domh", true, &cfg->domh))) return rc;
    if ((rc = get_double(p, doc, "grid", "span", true, &cfg->span))) return rc;
    if ((rc = get_double(p, doc, "physics", "amach1", true, &cfg->amach.x))) return rc;
    if ((rc = get_double(p, doc, "physics", "amach2", true, &cfg->amach.y))) return rc;
    if ((rc = get_double(p, doc, "physics", "amach3", true, &cfg->amach.z))) return rc;
    if (cfg->doml0 + cfg->doml1 == 0.0 || cfg->domh == 0.0 || cfg->span == 0.0)
        return fail(CS_ERR_ARG, p + ": the grid spans (doml0 + doml1, domh, span) must be non-zero");
    // optional run section (constants of driver/canard.cu:117-127 by default)
    if ((rc = get_double(p, doc, "run", "cfl", false, &cfg->cfl))) return rc;
    if ((rc = get_double(p, doc, "run", "tmax", false, &cfg->tmax))) return rc;
    if ((rc = get_int(p, doc, "run", "ndata", false, &cfg->ndata))) return rc;
    if ((rc = get_int(p, doc, "run", "nsteps", false, &cfg->nsteps))) return rc;
    if ((rc = get_int(p, doc, "run", "viscous", false, &cfg->viscous))) return rc;
    if ((rc = get_double(p, doc, "run", "re", false, &cfg->re))) return rc;
    if ((rc = get_double(p, doc, "run", "alpha_f", false, &cfg->alpha_f))) return rc;
    if ((rc = get_int(p, doc, "run", "filter", false, &cfg->filter))) return rc;
    if ((rc = get_double(p, doc, "run", "pulse_amp", false, &cfg->pulse_amp))) return rc;
    if ((rc = get_double(p, doc, "run", "pulse_width", false, &cfg->pulse_width))) return rc;
    if (!(cfg->cfl > 0.0) || cfg->ndata < 1 || cfg->nsteps < 0 || !(cfg->re > 0.0) || !(cfg->pulse_width > 0.0))
        return fail(CS_ERR_ARG, p + ": run section out of range (cfl > 0, ndata >= 1, nsteps >= 0, re > 0, "
                                    "pulse_width > 0)");
    return CS_OK;
}

extern "C" int cs_config_dims(const cs_run_config* cfg, cs_dims* dims, int* nbc6) {
    if (!cfg || !dims) return fail(CS_ERR_ARG, "null argument");
    dims->lxi = cfg->lximb + 1;
    dims->let = cfg->letmb + 1;
    dims->lze = cfg->lzemb + 1;
    if (nbc6)
        for (int i = 0; i < 6; i++) nbc6[i] = cfg->nbbc[i];
    return CS_OK;
}

extern "C" int cs_config_grid(const cs_run_config* cfg, double* xyz) {
    if (!cfg || !xyz) return fail(CS_ERR_ARG, "null argument");
    const int lxi0 = cfg->lximb, let0 = cfg->letmb, lze0 = cfg->lzemb;
    const long np = (long)(lxi0 + 1) * (let0 + 1) * (lze0 + 1);
    // grid.hpp:89-110: x from -doml0 to doml1, y from -domh to domh, z from
    // span/2 down to -span/2
    const double ra0 = (cfg->doml1 + cfg->doml0) / lxi0;
    const double ra1 = 2.0 * cfg->domh / let0;
    for (int k = 0; k <= lze0; k++)
        for (int j = 0; j <= let0; j++)
            for (int i = 0; i <= lxi0; i++) {
                const long l = ((long)k * (let0 + 1) + j) * (lxi0 + 1) + i;
                xyz[l] = -cfg->doml0 + i * ra0;
                xyz[np + l] = -cfg->domh + j * ra1;
                xyz[2 * np + l] = cfg->span * ((double)(lze0 - k) / lze0 - 0.5);
            }
    return CS_OK;
}

extern "C" int cs_vtk_write(const char* path, const cs_dims* dims, const double* xyz, int nvars,
                            const char* const* names, const double* const* vars, int precision) {
    if (!path || !dims || !xyz || (nvars > 0 && (!names || !vars))) return fail(CS_ERR_ARG, "null argument");
    if (dims->lxi < 1 || dims->let < 1 || dims->lze < 1) return fail(CS_ERR_SHAPE, "dimensions must be positive");
    if (nvars < 0) return fail(CS_ERR_ARG, "nvars must be >= 0");
    if (precision != 32 && precision != 64) return fail(CS_ERR_ARG, "precision must be 32 or 64");
    for (int v = 0; v < nvars; v++) {
        if (!names[v] || !vars[v]) return fail(CS_ERR_ARG, "null variable name or field");
        const std::string nm(names[v]);
        if (nm.empty() || nm.find_first_of(" \t\r\n") != std::string::npos)
            return fail(CS_ERR_ARG, "variable names must be non-empty words: '" + nm + "'");
    }
    std::ofstream out(path, std::ios::out | std::ios::trunc | std::ios::binary);
    if (!out) return fail(CS_ERR_ARG, std::string("cannot open ") + path + " for writing");
    const long np = (long)dims->lxi * dims->let * dims->lze;
    const char* ty = precision == 32 ? "float" : "double";
    // IO.hpp:117-153: legacy VTK, BINARY (big-endian), STRUCTURED_GRID, the
    // points interleaved x, y, z, then one SCALARS block per variable
    out << "# vtk DataFile Version 2.0\n";
    out << "Output compact schemes\n";
    out << "BINARY\n";
    out << "DATASET STRUCTURED_GRID\n";
    out << "DIMENSIONS " << dims->lxi << " " << dims->let << " " << dims->lze << "\n";
    out << "POINTS " << np << " " << ty << "\n";
    for (long i = 0; i < np; i++)
        for (int c = 0; c < 3; c++) {
            if (precision == 32) put_be32(out, (float)xyz[c * np + i]);
            else put_be64(out, xyz[c * np + i]);
        }
    if (nvars > 0) out << "\nPOINT_DATA " << np << "\n";
    for (int v = 0; v < nvars; v++) {
        out << "SCALARS " << names[v] << " " << ty << " 1\n";
        out << "LOOKUP_TABLE default\n";
        for (long i = 0; i < np; i++) {
            if (precision == 32) put_be32(out, (float)vars[v][i]);
            else put_be64(out, vars[v][i]);
        }
        out << "\n";
    }
    out.close();
    if (!out) return fail(CS_ERR_ARG, std::string("write to ") + path + " failed");
    return CS_OK;
}
