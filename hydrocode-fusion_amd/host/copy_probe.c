/* copy_probe.out -- streams one device buffer into another through the
 * library's plain copy kernel (hgrp_copy_probe, copy_probe_kernel) and prints
 * the bytes moved per copy and the device time.  It is the known-size kernel
 * against which the HBM byte counters (rocprofv3 --pmc FETCH_SIZE /
 * WRITE_SIZE) are calibrated: each copy reads and writes exactly `bytes`.
 *
 *   copy_probe.out [MiB per copy, default 1024] [copies, default 4]
 */
#include <stdio.h>
#include <stdlib.h>

#include "hydrocode_grp.h"

int main(int argc, char **argv)
{
    const long mib = argc > 1 ? atol(argv[1]) : 1024;
    const int reps = argc > 2 ? atoi(argv[2]) : 4;
    if (mib < 1 || mib > 65536 || reps < 1 || reps > 1000) {
        fprintf(stderr, "usage: %s [MiB per copy, 1..65536] [copies, 1..1000]\n", argv[0]);
        return 2;
    }
    const size_t bytes = (size_t)mib << 20;
    double ms = 0.0;
    const int rc = hgrp_copy_probe(bytes, reps, &ms);
    if (rc != 0) {
        fprintf(stderr, "copy_probe: hgrp_copy_probe failed (%d)\n", rc);
        return 1;
    }
    printf("copy of %zu bytes (%ld MiB, %zu KiB read + %zu KiB written) x %d: %.4f ms per copy, %.1f GB/s\n",
           bytes, mib, bytes >> 10, bytes >> 10, reps, ms, 2.0 * (double)bytes / (ms * 1e-3) / 1e9);
    return 0;
}
