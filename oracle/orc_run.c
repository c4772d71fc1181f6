/*
 * Test infrastructure: command-line driver of the C oracle.
 *   orc_run CASE OUT_DIR            run and write the requested dumps
 *   orc_run CASE OUT_DIR --time N   time N steps, print seconds, dump last
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "biocfd_oracle.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: orc_run CASE OUT_DIR [--time NSTEPS]\n");
        return 2;
    }
    if (argc >= 5 && strcmp(argv[3], "--time") == 0) {
        double sec = 0.0;
        int64_t n = atoll(argv[4]);
        int rc = orc_run_case(argv[1], argv[2], n, &sec);
        if (rc) { fprintf(stderr, "orc_run: error %d\n", rc); return 1; }
        printf("steps %lld seconds %.9f\n", (long long)n, sec);
        return 0;
    }
    int rc = orc_run_case(argv[1], argv[2], 0, NULL);
    if (rc) { fprintf(stderr, "orc_run: error %d\n", rc); return 1; }
    return 0;
}
