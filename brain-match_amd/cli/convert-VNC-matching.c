/* Drop-in for convert-VNC-matching (src/convert-VNC-matching.c): strips the
 * 'm' and 'f' prefixes from a matching file.
 * Usage: convert-VNC-matching <input_matching_file.csv>
 * Writes <input minus its extension>_generic.csv (:34-44); exit status 1 on
 * usage, file or line errors. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "brainmatch.h"

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "Usage: %s <input_matching_file.csv>\n", argv[0]);
        fprintf(stderr, "Strips 'm' and 'f' prefixes from node IDs and creates generic_matching.csv\n");
        return 1;
    }
    const char* dot = strrchr(argv[1], '.');
    const size_t stem = dot ? (size_t)(dot - argv[1]) : strlen(argv[1]);
    char* out = (char*)malloc(stem + 16);
    if (!out) return 1;
    memcpy(out, argv[1], stem);
    strcpy(out + stem, "_generic.csv");
    int64_t lines = 0;
    if (bm_convert_mapping_csv(argv[1], out, 0, &lines) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        free(out);
        return 1;
    }
    printf("Successfully processed %lld lines\n", (long long)lines);
    printf("Created %s\n", out);
    free(out);
    return 0;
}
