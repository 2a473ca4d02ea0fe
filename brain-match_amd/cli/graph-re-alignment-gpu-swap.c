/* Drop-in for src/graph-re-alignment-gpu-swap.cu: score-weighted 100-101 pair
 * perturbations, 2-swap candidates. */
#include "bm_cli.h"

int main(int argc, char** argv) {
    return bm_tool_main(argc, argv, BM_SEARCH_SWAP_WEIGHTED, "Graph Alignment Tool (MI355X swap deltas)");
}
