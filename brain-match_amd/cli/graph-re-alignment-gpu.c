/* Drop-in for src/graph-re-alignment-gpu.cu: uniform 3-5 pair perturbations,
 * 2-swap candidates. */
#include "bm_cli.h"

int main(int argc, char** argv) {
    return bm_tool_main(argc, argv, BM_SEARCH_SWAP_UNIFORM, "Graph Alignment Tool (MI355X swap deltas, uniform proposals)");
}
