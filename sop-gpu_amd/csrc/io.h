// io.h -- file formats of the reference: .top topologies, PDB/XYZ
// coordinates, PDB and DCD output.
#pragma once
#include <cstdio>
#include <string>
#include <vector>

namespace sop {

struct Atom {  // PDBAtom, src/IO/pdbio.h
    int id = 0;
    std::string name, resName;
    char chain = ' ', altLoc = ' ';
    int resid = 0;
    double x = 0, y = 0, z = 0, occupancy = 0, beta = 0;
};

struct Topology {  // SOP, src/IO/topio.h
    std::vector<Atom> atoms;
    std::vector<int> bi, bj;
    std::vector<float> br0;
    std::vector<int> ni, nj;
    std::vector<float> nr0, neh;
    std::vector<int> pi, pj;
};

// SOP::load (src/IO/topio.cpp:36-80); throws ConfigError on a malformed file.
Topology read_topology(const std::string& path);
// readCoord (src/IO/pdbio.cpp:292-337): fills atoms[].x/y/z from ATOM
// records of a PDB, or from a .xyz file (two header lines, `name x y z`).
void read_coordinates(const std::string& path, Topology& top);
// savePDB (src/IO/pdbio.cpp:194-219): ATOM records (hexadecimal serials
// from 100000 atoms on), CONECT records of the bonds, END.
void write_pdb(const std::string& path, const Topology& top, const std::vector<Atom>& extra = {});

// DCD writer (src/IO/dcdio.cpp:100-194): CHARMM/NAMD header and frames.
struct DcdWriter {
    static void header(const std::string& path, int natoms, int nfile, int npriv, int nsavc, float delta);
    static void frame(const std::string& path, const std::vector<float>& x, const std::vector<float>& y,
                      const std::vector<float>& z);
};

}  // namespace sop
