// ref_harness.cpp -- drives the reference's own host sources (built from
// BarsegovGroup/SOP-GPU src/ where they lie, see oracle/Makefile) to emit
// the vectors that pin the oracle and the engine's readers/writers:
//
//   ref_harness seeds <seed> <n>          seed table of ht.cu:40-56 via the
//                                         reference ran2 (ran2.cpp:51-89)
//   ref_harness ran2 <idum> <count>       raw ran2 stream
//   ref_harness top <file.top>            SOP::load (topio.cpp:36-80) dump
//   ref_harness coords <top> <pdb|xyz>    readCoord (pdbio.cpp:292-337) dump
//   ref_harness savepdb <top> <coords> <out.pdb>   savePDB (pdbio.cpp:194-219)
//   ref_harness dcd <out.dcd> <N> <frames> DCD header + frames (dcdio.cpp)
//
// Test infrastructure only; the seeds loop below restates ht.cu:40-56 (a
// CUDA file that cannot be compiled here) around the reference ran2.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "IO/dcdio.h"
#include "IO/pdbio.h"
#include "IO/topio.h"
#include "Integrators/ran2.h"

static int cmd_seeds(int seed, int n) {
    for (int i = 0; i < n; i++) {
        unsigned v[4];
        for (int c = 0; c < 4; c++) {
            do {
                v[c] = (unsigned)(ran2(&seed) * UINT_MAX);
            } while (v[c] < 128);
        }
        printf("%u %u %u %u\n", v[0], v[1], v[2], v[3]);
    }
    return 0;
}

static int cmd_ran2(int idum, int count) {
    for (int i = 0; i < count; i++) printf("%.17g %d\n", ran2(&idum), idum);
    return 0;
}

static int cmd_top(const char* path) {
    SOP sop;
    sop.load(path);
    printf("atoms %zu\n", sop.aminos.size());
    for (size_t i = 0; i < sop.aminos.size(); i++) {
        const PDBAtom& a = sop.aminos[i];
        printf("A %d %s %d %s %c %.6f %.6f\n", a.id, a.name, a.resid, a.resName, a.chain, a.occupancy, a.beta);
    }
    printf("bonds %zu\n", sop.bonds.size());
    for (size_t i = 0; i < sop.bonds.size(); i++)
        printf("B %d %d %.9g\n", sop.bonds[i].i, sop.bonds[i].j, sop.bonds[i].r0);
    printf("natives %zu\n", sop.natives.size());
    for (size_t i = 0; i < sop.natives.size(); i++)
        printf("N %d %d %.9g %.9g\n", sop.natives[i].i, sop.natives[i].j, sop.natives[i].r0, sop.natives[i].eh);
    printf("pairs %zu\n", sop.pairs.size());
    for (size_t i = 0; i < sop.pairs.size(); i++) printf("P %d %d\n", sop.pairs[i].i, sop.pairs[i].j);
    return 0;
}

static int cmd_coords(const char* top, const char* crd) {
    SOP sop;
    sop.load(top);
    readCoord(crd, sop);
    for (size_t i = 0; i < sop.aminos.size(); i++) {
        // gsop.cu:185-199 stores them as float
        printf("%.9g %.9g %.9g\n", (float)sop.aminos[i].x, (float)sop.aminos[i].y, (float)sop.aminos[i].z);
    }
    return 0;
}

static int cmd_savepdb(const char* top, const char* crd, const char* out) {
    SOP sop;
    sop.load(top);
    readCoord(crd, sop);
    for (size_t i = 0; i < sop.aminos.size(); i++) {
        // the run writes float coordinates back into the SOP (gsop.cu:155-158)
        sop.aminos[i].x = (float)sop.aminos[i].x;
        sop.aminos[i].y = (float)sop.aminos[i].y;
        sop.aminos[i].z = (float)sop.aminos[i].z;
    }
    savePDB(out, sop);
    return 0;
}

static int cmd_dcd(const char* out, int N, int frames) {
    DCD dcd;
    dcd.open_write(out);
    dcd.N = N;
    dcd.NFILE = dcd.NPRIV = dcd.NSAVC = 1000;
    dcd.DELTA = 0.0;
    dcd.write_header();
    dcd.close();
    float* X = (float*)malloc(sizeof(float) * N);
    float* Y = (float*)malloc(sizeof(float) * N);
    float* Z = (float*)malloc(sizeof(float) * N);
    for (int f = 0; f < frames; f++) {
        for (int i = 0; i < N; i++) {
            X[i] = 0.5f * i + f;
            Y[i] = -0.25f * i + 2.0f * f;
            Z[i] = 1.0f / (1.0f + i) - f;
        }
        dcd.open_append(out);
        dcd.write_frame(X, Y, Z);
        dcd.close();
    }
    free(X); free(Y); free(Z);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: ref_harness seeds|ran2|top|coords|savepdb|dcd ...\n");
        return 2;
    }
    // the reference also logs to stdout; dump lines carry a tag so readers
    // can tell them apart
    const char* c = argv[1];
    if (!strcmp(c, "seeds") && argc == 4) return cmd_seeds(atoi(argv[2]), atoi(argv[3]));
    if (!strcmp(c, "ran2") && argc == 4) return cmd_ran2(atoi(argv[2]), atoi(argv[3]));
    if (!strcmp(c, "top") && argc == 3) return cmd_top(argv[2]);
    if (!strcmp(c, "coords") && argc == 4) return cmd_coords(argv[2], argv[3]);
    if (!strcmp(c, "savepdb") && argc == 5) return cmd_savepdb(argv[2], argv[3], argv[4]);
    if (!strcmp(c, "dcd") && argc == 5) return cmd_dcd(argv[2], atoi(argv[3]), atoi(argv[4]));
    fprintf(stderr, "bad arguments\n");
    return 2;
}
