/* Drop-in for src/graph-re-alignment-gpu-3-swap.cu: uniform 3-5 pair
 * perturbations, 3-vertex permutation candidates. */
#include "bm_cli.h"

int main(int argc, char** argv) {
    return bm_tool_main(argc, argv, BM_SEARCH_ROT3, "Graph Alignment Tool (MI355X 3-vertex permutation deltas)");
}
