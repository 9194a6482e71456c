/*
 * triv_par_mpi.c -- the oracle's scrimp_triv_par as an MPI program, one
 * process per rank, the way the reference runs (`mpirun -n <cores> scrimppp
 * scrimp_triv_par ...`).  TEST AND BASELINE INFRASTRUCTURE ONLY: it times the
 * CPU baseline; the product library never links it.
 *
 *   mpirun -n R triv_par_mpi <series.f64> <m> [block_stride] [blocklen] [kernel]
 *
 * <series.f64> holds raw little-endian doubles.  Every rank reads the whole
 * series (ScrimpTrivialPar.cpp:142-160 via fetch_time_series), computes its
 * share with orc_triv_par_rank (statistics, partitions r and 2R-1-r, blocks of
 * blocklen = 500 with their seeds, the AVX2 kernel, the d2 transform;
 * ScrimpTrivialPar.cpp:201-340) and the partial profiles are merged on rank 0
 * by MPI_Reduce with a user-defined operation that takes an element when it
 * is strictly smaller (ScrimpTrivialPar.cpp:68-108, 186-198, 343-358).
 * block_stride > 1 evaluates only every block_stride-th block of each rank (a
 * bounded timing sample).  Timed like the reference with IGNORE_INIT_TIME
 * (main.cpp:84-90): from a barrier after reading the input to the end of the
 * reduction, i.e. the slowest rank's statistics + seeds + evaluation plus the
 * reduction.  Rank 0 prints one JSON line.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scrimp_oracle.h"

static long g_P = 0;

/* element-wise min of (d2, index) SoAs laid out as [P doubles][P longs] */
static void soa_min(void* in, void* inout, int* len, MPI_Datatype* dt) {
    (void)dt;
    const double* ip = (const double*)in;
    const long* ii = (const long*)(ip + g_P);
    double* op = (double*)inout;
    long* oi = (long*)(op + g_P);
    for (int r = 0; r < *len; ++r) {
        for (long i = 0; i < g_P; ++i) {
            if (ip[i] < op[i]) {
                op[i] = ip[i];
                oi[i] = ii[i];
            }
        }
    }
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 3) {
        if (rank == 0) fprintf(stderr, "usage: %s <series.f64> <m> [block_stride] [blocklen] [kernel]\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    const int m = atoi(argv[2]);
    const int stride = argc > 3 ? atoi(argv[3]) : 1;
    const int blocklen = argc > 4 ? atoi(argv[4]) : 500;
    const int kernel = argc > 5 ? atoi(argv[5]) : 1;
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "rank %d: cannot open %s\n", rank, argv[1]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f) / (long)sizeof(double);
    fseek(f, 0, SEEK_SET);
    /* 64-byte aligned like the reference's aligned vectors (ScrimpTrivialPar.cpp:37-56) */
    double* A = NULL;
    if (posix_memalign((void**)&A, 64, sizeof(double) * (size_t)(n + 1)) != 0) A = NULL;
    if (!A || fread(A, sizeof(double), (size_t)n, f) != (size_t)n) {
        fprintf(stderr, "rank %d: cannot read %s\n", rank, argv[1]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    fclose(f);
    A[n] = 0.0; /* the kernel's last update reads one element past the series */
    const long P = n - m + 1;
    if (P <= 0 || m < 1) {
        if (rank == 0) fprintf(stderr, "parameters chosen badly: resulting matrix profile of length 0\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    g_P = P;
    char* soa = NULL;
    if (posix_memalign((void**)&soa, 64, (size_t)P * (sizeof(double) + sizeof(long))) != 0) MPI_Abort(MPI_COMM_WORLD, 2);
    double* prof = (double*)soa;
    long* idx = (long*)(prof + P);
    MPI_Datatype soa_type;
    MPI_Type_contiguous((int)(P * (long)(sizeof(double) + sizeof(long))), MPI_BYTE, &soa_type);
    MPI_Type_commit(&soa_type);
    MPI_Op op;
    MPI_Op_create(soa_min, 1, &op);

    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    double cells = 0.0, secs = 0.0;
    if (orc_triv_par_rank(A, n, m, -1, blocklen, rank, size, kernel, stride, prof, idx, &cells, &secs) != 0)
        MPI_Abort(MPI_COMM_WORLD, 3);
    if (rank == 0)
        MPI_Reduce(MPI_IN_PLACE, soa, 1, soa_type, op, 0, MPI_COMM_WORLD);
    else
        MPI_Reduce(soa, NULL, 1, soa_type, op, 0, MPI_COMM_WORLD);
    const double t_total = MPI_Wtime() - t0;
    double cells_sum = 0.0, secs_max = 0.0, total_max = 0.0;
    MPI_Reduce(&cells, &cells_sum, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    MPI_Reduce(&secs, &secs_max, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    MPI_Reduce(&t_total, &total_max, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        double checksum = 0.0;
        long valid = 0;
        for (long i = 0; i < P; ++i) {
            if (idx[i] >= 0) {
                checksum += prof[i];
                ++valid;
            }
        }
        printf("{\"ranks\": %d, \"n\": %ld, \"m\": %d, \"block_stride\": %d, \"blocklen\": %d, \"cells\": %.17g, "
               "\"seconds\": %.9g, \"slowest_rank_compute_s\": %.9g, \"entries_with_partner\": %ld, "
               "\"checksum\": %.17g}\n",
               size, n, m, stride, blocklen, cells_sum, total_max, secs_max, valid, checksum);
        fflush(stdout);
    }
    MPI_Op_free(&op);
    MPI_Type_free(&soa_type);
    free(soa);
    free(A);
    MPI_Finalize();
    return 0;
}
