/* Drop-in for transform-solution (src/transform-solution.c): adds the 'm'
 * and 'f' prefixes to a "male,female" matching file.
 * Usage: transform-solution <input_solution_file.csv>
 * Writes <input minus its extension>_matching.csv (:42-52); exit status 1 on
 * usage, file or line errors (:30-33, :37-40, :73-89). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "brainmatch.h"

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "Usage: %s <input_solution_file.csv>\n", argv[0]);
        fprintf(stderr, "Converts solution file to matching format by adding 'm' and 'f' prefixes\n");
        return 1;
    }
    const char* dot = strrchr(argv[1], '.');
    const size_t stem = dot ? (size_t)(dot - argv[1]) : strlen(argv[1]);
    char* out = (char*)malloc(stem + 16);
    if (!out) return 1;
    memcpy(out, argv[1], stem);
    strcpy(out + stem, "_matching.csv");
    int64_t lines = 0;
    if (bm_convert_mapping_csv(argv[1], out, 1, &lines) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        free(out);
        return 1;
    }
    printf("Successfully processed %lld lines\n", (long long)lines);
    printf("Created %s\n", out);
    free(out);
    return 0;
}
