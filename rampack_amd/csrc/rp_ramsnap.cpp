// RAMSNAP snapshots (rampack's text snapshot format, docs/output-formats.md): host-side reader and writer behind
// the C ABI.  Layout: the number of auxiliary entries, one "key value" line each (sorted by key), the 9 entries of
// the box matrix on one line (a line of 3 numbers is the older cuboid form, read as a diagonal box), the number of
// particles, then per particle its position and its row-major orientation.  Numbers are written and parsed by the
// C++ standard streams with 17 significant digits, which round-trips every double.
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "rampack_amd.h"

namespace rp {
int set_error(int code, const std::string &msg);  // rp_capi.cpp
}

namespace {

using AuxMap = std::map<std::string, std::string>;

// "key value" lines -> map; the key is the first whitespace-free token, the value the rest of the line after
// leading whitespace (RamsnapReader::restoreAuxInfo, src/core/io/RamsnapReader.cpp:48-63)
void parseAux(const char *text, AuxMap &out) {
    if (!text) return;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key, value;
        if (!(ls >> key)) continue;  // blank line
        ls >> std::ws;
        std::getline(ls, value);
        out[key] = value;
    }
}

// RamsnapWriter::write (src/core/io/RamsnapWriter.cpp:8-48)
std::string format(const double *box, size_t n, const double *pos, const double *rot, const AuxMap &aux) {
    std::ostringstream out;
    out.precision(std::numeric_limits<double>::max_digits10);
    out << std::defaultfloat;
    out << aux.size() << std::endl;
    for (const auto &kv : aux) out << kv.first << " " << kv.second << std::endl;
    out << box[0] << " " << box[1] << " " << box[2] << " ";
    out << box[3] << " " << box[4] << " " << box[5] << " ";
    out << box[6] << " " << box[7] << " " << box[8] << std::endl;
    out << n << std::endl;
    for (size_t i = 0; i < n; i++) {
        const double *p = pos + 3 * i, *o = rot + 9 * i;
        out << p[0] << " " << p[1] << " " << p[2];
        out << "        ";
        out << o[0] << " " << o[1] << " " << o[2] << " ";
        out << o[3] << " " << o[4] << " " << o[5] << " ";
        out << o[6] << " " << o[7] << " " << o[8];
        out << std::endl;
    }
    return out.str();
}

struct Snapshot {
    AuxMap aux;
    double box[9] = {};
    std::vector<double> pos, rot;
};

double det3(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// RamsnapReader::read (src/core/io/RamsnapReader.cpp:15-92)
int parse(std::istream &in, Snapshot &s, bool shapes) {
    auto broken = [](const std::string &what) { return rp::set_error(RP_ERR_VALIDATION, "Broken RAMSNAP file: " + what); };
    size_t nAux = 0;
    if (!(in >> nAux)) return broken("aux info size");
    // with no entries the box line follows the count line; the reference's reader starts the box line at the end
    // of the count line instead and so rejects such files, including those its own writer produces
    if (nAux == 0) in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    for (size_t i = 0; i < nAux; i++) {
        std::string key, value;
        in >> key >> std::ws;
        std::getline(in, value);
        if (!in) return broken("aux info entry " + std::to_string(i));
        s.aux[key] = value;
    }
    std::string line;
    std::getline(in, line);
    if (!in) return broken("dimensions");
    std::istringstream ds(line);
    double t[3];
    if (!(ds >> t[0] >> t[1] >> t[2])) return broken("dimensions");
    ds >> std::ws;
    if (!ds.good()) {
        // three numbers: the older cuboid form
        s.box[0] = t[0];
        s.box[4] = t[1];
        s.box[8] = t[2];
    } else {
        s.box[0] = t[0];
        s.box[1] = t[1];
        s.box[2] = t[2];
        if (!(ds >> s.box[3] >> s.box[4] >> s.box[5] >> s.box[6] >> s.box[7] >> s.box[8])) return broken("dimensions");
    }
    if (det3(s.box) == 0) return broken("zero box volume");
    size_t n = 0;
    if (!(in >> n)) return broken("size");
    if (!shapes) {
        s.pos.resize(3 * n);  // size only
        return RP_OK;
    }
    s.pos.assign(3 * n, 0.0);
    s.rot.assign(9 * n, 0.0);
    for (size_t i = 0; i < n; i++) {
        double *p = &s.pos[3 * i], *o = &s.rot[9 * i];
        in >> p[0] >> p[1] >> p[2];
        for (int c = 0; c < 9; c++) in >> o[c];
        if (!in) return broken("shape " + std::to_string(i) + "/" + std::to_string(n));
    }
    return RP_OK;
}

std::string auxText(const AuxMap &aux) {
    std::string t;
    for (const auto &kv : aux) t += kv.first + " " + kv.second + "\n";
    return t;
}

}  // namespace

extern "C" {

int rp_ramsnap_write(const char *path, const double box[9], size_t n, const double *pos, const double *rot,
                     const char *aux_text) {
    if (!path || !box || (n && (!pos || !rot))) return rp::set_error(RP_ERR_PRECONDITION, "null argument");
    AuxMap aux;
    parseAux(aux_text, aux);
    std::ofstream f(path, std::ios::binary);
    if (!f) return rp::set_error(RP_ERR_PRECONDITION, std::string("cannot open ") + path + " for writing");
    f << format(box, n, pos, rot, aux);
    f.close();
    if (!f) return rp::set_error(RP_ERR_PRECONDITION, std::string("cannot write ") + path);
    return RP_OK;
}

int rp_ramsnap_info(const char *path, size_t *n, size_t *aux_bytes) {
    if (!path || !n || !aux_bytes) return rp::set_error(RP_ERR_PRECONDITION, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return rp::set_error(RP_ERR_PRECONDITION, std::string("cannot open ") + path);
    Snapshot s;
    int r = parse(f, s, false);
    if (r) return r;
    *n = s.pos.size() / 3;
    *aux_bytes = auxText(s.aux).size() + 1;
    return RP_OK;
}

int rp_ramsnap_read(const char *path, double box[9], size_t n, double *pos, double *rot, char *aux_text,
                    size_t aux_cap) {
    if (!path || !box || (n && (!pos || !rot))) return rp::set_error(RP_ERR_PRECONDITION, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return rp::set_error(RP_ERR_PRECONDITION, std::string("cannot open ") + path);
    Snapshot s;
    int r = parse(f, s, true);
    if (r) return r;
    if (s.pos.size() / 3 != n)
        return rp::set_error(RP_ERR_PRECONDITION, "the file holds " + std::to_string(s.pos.size() / 3) +
                                                      " particles, the buffers " + std::to_string(n));
    std::memcpy(box, s.box, sizeof s.box);
    if (n) {
        std::memcpy(pos, s.pos.data(), sizeof(double) * 3 * n);
        std::memcpy(rot, s.rot.data(), sizeof(double) * 9 * n);
    }
    if (aux_text) {
        std::string t = auxText(s.aux);
        if (t.size() + 1 > aux_cap) return rp::set_error(RP_ERR_PRECONDITION, "aux buffer too small");
        std::memcpy(aux_text, t.c_str(), t.size() + 1);
    }
    return RP_OK;
}

}  // extern "C"
