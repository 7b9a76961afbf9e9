// io.cpp -- .top / PDB / XYZ readers and PDB / DCD writers with the
// reference's field layouts (src/IO/topio.cpp, pdbio.cpp, dcdio.cpp).
#include "io.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "../../include/sopgpu.h"
#include "config.h"
#include "hostutil.h"

namespace sop {

namespace {

enum Section { NONE, ATOMS, BONDS, NATIVES, PAIRS };

Section section_of(const char* line) {
    if (std::strncmp(line, "[ atoms ]", 9) == 0) return ATOMS;
    if (std::strncmp(line, "[ bonds ]", 9) == 0) return BONDS;
    if (std::strncmp(line, "[ native ]", 10) == 0) return NATIVES;
    if (std::strncmp(line, "[ pairs ]", 9) == 0) return PAIRS;
    return NONE;
}

// Space-separated fields, as strtok(line, " ") splits them.
std::vector<std::string> fields(const char* line) {
    std::vector<std::string> f;
    std::string cur;
    for (const char* p = line; *p; ++p) {
        if (*p == ' ') {
            if (!cur.empty()) f.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(*p);
        }
    }
    if (!cur.empty()) f.push_back(cur);
    return f;
}

}  // namespace

Topology read_topology(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw ConfigError("Can't open topology file '" + path + "'");
    Topology t;
    char buf[256];
    Section sec = NONE;
    int line = 0;
    while (std::fgets(buf, sizeof buf, f)) {
        ++line;
        if (buf[0] == ';' || buf[0] == '\n' || buf[0] == '\0') continue;
        if (buf[0] == '[') {
            sec = section_of(buf);
            if (sec == NONE) {
                std::fclose(f);
                throw ConfigError("Malformed section header at " + path + ":" + std::to_string(line));
            }
            continue;
        }
        const std::vector<std::string> v = fields(buf);
        switch (sec) {
        case ATOMS: {
            if (v.size() < 8) {
                std::fclose(f);
                throw ConfigError("Malformed atom at " + path + ":" + std::to_string(line));
            }
            Atom a;
            a.id = std::atoi(v[0].c_str());
            a.resid = std::atoi(v[2].c_str());
            a.resName = v[3];
            a.name = v[4];
            a.chain = v[5][0];
            a.occupancy = std::atof(v[6].c_str());
            a.beta = std::atof(v[7].c_str());
            t.atoms.push_back(a);
            break;
        }
        case BONDS:
        case NATIVES:
        case PAIRS: {
            if (v.size() < 3) {
                std::fclose(f);
                throw ConfigError("Malformed pair at " + path + ":" + std::to_string(line));
            }
            const int i = std::atoi(v[0].c_str()), j = std::atoi(v[1].c_str());
            const float c0 = v.size() > 3 ? (float)std::atof(v[3].c_str()) : 0.0f;
            const float c1 = v.size() > 4 ? (float)std::atof(v[4].c_str()) : 0.0f;
            if (sec == BONDS) {
                t.bi.push_back(i);
                t.bj.push_back(j);
                t.br0.push_back(c0);
            } else if (sec == NATIVES) {
                t.ni.push_back(i);
                t.nj.push_back(j);
                t.nr0.push_back(c0);
                t.neh.push_back(c1);
            } else {
                t.pi.push_back(i);
                t.pj.push_back(j);
            }
            break;
        }
        case NONE:
            std::fclose(f);
            throw ConfigError("Malformed topology: meaningful line at " + path + ":" + std::to_string(line) +
                              " outside any section");
        }
    }
    std::fclose(f);
    return t;
}

void read_coordinates(const std::string& path, Topology& top) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw ConfigError("ERROR: Coordinates file " + path + " can not be found.");
    const bool xyz = path.size() >= 4 && path.compare(path.size() - 4, 4, ".xyz") == 0;
    char buf[81 + 1];
    size_t index = 0;
    if (xyz) {
        for (int k = 0; k < 2; ++k)
            if (!std::fgets(buf, 80, f)) break;
    }
    while (std::fgets(buf, 80, f)) {
        if (xyz) {
            char* p = std::strtok(buf, " \t\r\n");
            double c[3] = {0, 0, 0};
            for (int d = 0; d < 3; ++d) {
                p = std::strtok(nullptr, " \t\r\n");
                c[d] = p ? std::atof(p) : 0.0;
            }
            if (index < top.atoms.size()) {
                top.atoms[index].x = c[0];
                top.atoms[index].y = c[1];
                top.atoms[index].z = c[2];
            }
            ++index;
        } else if (std::strncmp(buf, "ATOM", 4) == 0) {
            auto col = [&](int a, int len) {
                char t[16];
                std::strncpy(t, buf + a, len);
                t[len] = '\0';
                return std::atof(t);
            };
            if (index < top.atoms.size()) {
                top.atoms[index].x = col(30, 8);
                top.atoms[index].y = col(38, 8);
                top.atoms[index].z = col(46, 8);
            }
            ++index;
        }
    }
    std::fclose(f);
    if (index == 0) throw ConfigError("Can't read pdb file.");
    if (index != top.atoms.size())
        throw ConfigError("Read coordinates for " + std::to_string(index) + " beads, yet topology has " +
                          std::to_string(top.atoms.size()) + " beads.");
}

static void print_atom(FILE* f, const Atom& a, int serial, bool hex) {
    std::fprintf(f, hex ? "ATOM  %5x %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                        : "ATOM  %5d %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n",
                 serial, a.name.c_str(), a.altLoc, a.resName.c_str(), a.chain, a.resid, a.x, a.y, a.z, a.occupancy,
                 a.beta);
}

void write_pdb(const std::string& path, const Topology& top, const std::vector<Atom>& extra) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw ConfigError("Can't write '" + path + "'");
    const bool hex = top.atoms.size() + extra.size() >= 100000;
    for (size_t i = 0; i < top.atoms.size(); ++i) print_atom(f, top.atoms[i], (int)i + 1, hex);
    for (size_t i = 0; i < extra.size(); ++i) print_atom(f, extra[i], (int)(top.atoms.size() + i + 1), hex);
    int last = -1;
    for (size_t k = 0; k < top.bi.size(); ++k) {
        if (last != top.bi[k]) {
            if (last != -1) std::fprintf(f, "\n");
            std::fprintf(f, "CONECT%5d", top.bi[k] + 1);
            last = top.bi[k];
        }
        std::fprintf(f, "%5d", top.bj[k] + 1);
    }
    std::fprintf(f, "\nEND");
    std::fclose(f);
}

static void put_i(FILE* f, int v) { std::fwrite(&v, 4, 1, f); }

void DcdWriter::header(const std::string& path, int natoms, int nfile, int npriv, int nsavc, float delta) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw ConfigError("Can't write '" + path + "'");
    put_i(f, 84);
    std::fwrite("CORD", 4, 1, f);
    put_i(f, nfile);
    put_i(f, npriv);
    put_i(f, nsavc);
    put_i(f, npriv - nsavc);
    for (int i = 0; i < 5; ++i) put_i(f, 0);
    std::fwrite(&delta, 4, 1, f);
    for (int i = 0; i < 9; ++i) put_i(f, 0);
    put_i(f, 24);
    put_i(f, 84);
    put_i(f, 164);
    put_i(f, 2);
    char title[81];
    std::memset(title, ' ', 80);
    title[80] = '\0';
    const char* r1 = "REMARKS CREATED BY dcdio.cpp";
    std::memcpy(title, r1, std::strlen(r1));
    std::fwrite(title, 80, 1, f);
    std::time_t now = std::time(nullptr);
    char date[128];
    std::snprintf(date, sizeof date, "REMARKS DATE: %s", std::asctime(std::localtime(&now)));
    std::memset(title, ' ', 80);
    std::memcpy(title, date, std::min<size_t>(80, std::strlen(date)));
    std::fwrite(title, 80, 1, f);
    put_i(f, 164);
    put_i(f, 4);
    put_i(f, natoms);
    put_i(f, 4);
    std::fclose(f);
}

void DcdWriter::frame(const std::string& path, const std::vector<float>& x, const std::vector<float>& y,
                      const std::vector<float>& z) {
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) throw ConfigError("Can't append to '" + path + "'");
    const int bytes = (int)x.size() * 4;
    for (const std::vector<float>* v : {&x, &y, &z}) {
        put_i(f, bytes);
        std::fwrite(v->data(), 4, v->size(), f);
        put_i(f, bytes);
    }
    std::fclose(f);
}

}  // namespace sop

namespace {

template <typename T>
T* dup_array(const std::vector<T>& v) {
    T* p = (T*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" int sop_read_topology(const char* top_path, const char* coords_path, sop_synth* out) {
    if (!top_path || !out) return sop::fail(SOP_E_ARG, "sop_read_topology: null argument");
    std::memset(out, 0, sizeof(*out));
    try {
        sop::Topology t = sop::read_topology(top_path);
        if (coords_path) sop::read_coordinates(coords_path, t);
        const int n = (int)t.atoms.size();
        std::vector<float> c4(4 * (size_t)n, 0.0f);
        for (int i = 0; i < n; ++i) {
            c4[4 * i] = (float)t.atoms[i].x;
            c4[4 * i + 1] = (float)t.atoms[i].y;
            c4[4 * i + 2] = (float)t.atoms[i].z;
        }
        out->n = n;
        out->coords4 = dup_array(c4);
        out->nbonds = (int)t.bi.size();
        out->bond_i = dup_array(t.bi);
        out->bond_j = dup_array(t.bj);
        out->bond_r0 = dup_array(t.br0);
        out->nnative = (int)t.ni.size();
        out->nat_i = dup_array(t.ni);
        out->nat_j = dup_array(t.nj);
        out->nat_r0 = dup_array(t.nr0);
        out->nat_eh = dup_array(t.neh);
        out->npairs = (int)t.pi.size();
        out->pair_i = dup_array(t.pi);
        out->pair_j = dup_array(t.pj);
        return 0;
    } catch (const sop::ConfigError& e) {
        return sop::fail(SOP_E_CONFIG, e.what());
    }
}

extern "C" int sop_write_pdb(const char* top_path, const float* coords4, int n, const char* out_path) {
    if (!top_path || !coords4 || !out_path) return sop::fail(SOP_E_ARG, "sop_write_pdb: null argument");
    try {
        sop::Topology t = sop::read_topology(top_path);
        if ((int)t.atoms.size() != n) return sop::fail(SOP_E_ARG, "sop_write_pdb: bead count differs from the topology");
        for (int i = 0; i < n; ++i) {
            t.atoms[i].x = coords4[4 * i];
            t.atoms[i].y = coords4[4 * i + 1];
            t.atoms[i].z = coords4[4 * i + 2];
        }
        sop::write_pdb(out_path, t);
        return 0;
    } catch (const sop::ConfigError& e) {
        return sop::fail(SOP_E_CONFIG, e.what());
    }
}

extern "C" int sop_write_dcd(const char* path, int natoms, int nframes, const float* frames, int freq, float delta) {
    if (!path || natoms < 1 || nframes < 0 || (nframes && !frames)) return sop::fail(SOP_E_ARG, "sop_write_dcd: bad arguments");
    try {
        sop::DcdWriter::header(path, natoms, freq, freq, freq, delta);
        for (int f = 0; f < nframes; ++f) {
            const float* b = frames + 3 * (size_t)natoms * f;
            std::vector<float> x(b, b + natoms), y(b + natoms, b + 2 * natoms), z(b + 2 * natoms, b + 3 * natoms);
            sop::DcdWriter::frame(path, x, y, z);
        }
        return 0;
    } catch (const sop::ConfigError& e) {
        return sop::fail(SOP_E_CONFIG, e.what());
    }
}
