// Dump tool linked against the reference's own host sources (Graph.cpp and
// QuickMapping.cpp from src/main, compiled in place by oracle/Makefile).  It
// lets the tests compare this repository's edgelist reader, CSR and pattern
// tables with what the reference code itself produces.  TEST INFRASTRUCTURE
// ONLY; nothing in the product links it.
//
//   ref_driver graph <edgelist> <out>   n, m, maxdeg, then one line per vertex
//                                       "v deg n1 n2 ..." (the CSR of
//                                       src/main/DuMatoCPU.cu:94-108 without
//                                       the -1 sentinel)
//   ref_driver dict <k> <exec_dir> <out>  numberOfQuicks, quickToCgLocal[],
//                                       numberOfCgs, cgLocalToGlobal[] as read
//                                       by QuickMapping(k) from
//                                       <exec_dir>/../dictionaries/k.csv
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <unistd.h>

#include "Graph.h"
#include "QuickMapping.h"

static int dump_graph(const char *path, const char *out) {
    FILE *chk = fopen(path, "r");
    if (!chk) return 2;
    fclose(chk);
    Graph g(path);
    FILE *fo = fopen(out, "w");
    if (!fo) return 3;
    int n = g.getNumberOfVertices();
    fprintf(fo, "%d %d %d\n", n, g.getNumberOfEdges(), g.getMaxDegree());
    for (int v = 0; v < n; v++) {
        std::set<int> *nb = g.getNeighbours(v);
        fprintf(fo, "%d %d", v, (int)nb->size());
        for (int u : *nb) fprintf(fo, " %d", u);
        fprintf(fo, "\n");
    }
    fclose(fo);
    return 0;
}

static int dump_dict(int k, const char *exec_dir, const char *out) {
    if (chdir(exec_dir) != 0) return 2;
    QuickMapping q(k);
    FILE *fo = fopen(out, "w");
    if (!fo) return 3;
    fprintf(fo, "%lu\n", q.numberOfQuicks);
    for (unsigned long i = 0; i < q.numberOfQuicks; i++)
        fprintf(fo, "%lu\n", q.quickToCgLocal[i]);
    fprintf(fo, "%lu\n", q.numberOfCgs);
    for (unsigned long i = 0; i < q.numberOfCgs; i++)
        fprintf(fo, "%lu\n", q.cgLocalToGlobal[i]);
    fclose(fo);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && !strcmp(argv[1], "graph")) return dump_graph(argv[2], argv[3]);
    if (argc == 5 && !strcmp(argv[1], "dict"))
        return dump_dict(atoi(argv[2]), argv[3], argv[4]);
    fprintf(stderr, "usage: ref_driver graph <edgelist> <out> | dict <k> <exec_dir> <out>\n");
    return 1;
}
