/* Linked into the timing build of the reference program (oracle/Makefile,
 * `make mpi`; test and benchmark infrastructure): two Fortran MPI bindings
 * in front of the library's own.
 *
 * mpi_reduce_ as an all-reduce: the result is defined on every rank, not only
 * on the root.  Why: src/computeOnm4.f90:211 tests
 * `iaq.eq.0.and.ihgntotal / iogntotal.ne.2` on every rank, and Fortran does
 * not require .and. to skip its second operand: the compiled code divides by
 * iogntotal on all ranks, where mpi_reduce (src/computeOnm4.f90:205) left it
 * undefined -- a rank whose stack slot happens to hold 0 takes SIGFPE.  The
 * reference's mpi_reduce calls (src/computeOnm4.f90:203-207,
 * src/get_molecule_count.f90:56-60) only feed diagnostic totals printed by
 * rank 0, so an all-reduce changes nothing the program computes.
 *
 * mpi_wtime_ that also logs: with SESMD_WTIME_LOG set to a directory, each
 * rank appends every value it returns to <dir>/wtime.<rank>, one line per
 * call, written at once (the program's own per-step print is buffered by the
 * Fortran runtime and lost when the run is stopped).  The program calls
 * mpi_wtime at the start of every step on every rank and at its end on rank 4
 * (src/mybox.f90:134,240). */
#include <mpi.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

void mpi_reduce_(void *sendbuf, void *recvbuf, MPI_Fint *count, MPI_Fint *datatype, MPI_Fint *op,
                 MPI_Fint *root, MPI_Fint *comm, MPI_Fint *ierr)
{
    (void)root;
    *ierr = MPI_Allreduce(sendbuf, recvbuf, (int)*count, MPI_Type_f2c(*datatype), MPI_Op_f2c(*op),
                          MPI_Comm_f2c(*comm));
}

static int log_fd = -2;  /* -2: not opened yet, -1: no log */

double mpi_wtime_(void)
{
    double t = MPI_Wtime();
    if (log_fd == -2) {
        const char *dir = getenv("SESMD_WTIME_LOG");
        log_fd = -1;
        if (dir) {
            int rank = 0;
            char path[4096];
            MPI_Comm_rank(MPI_COMM_WORLD, &rank);
            snprintf(path, sizeof path, "%s/wtime.%d", dir, rank);
            log_fd = open(path, O_WRONLY | O_CREAT | O_APPEND, 0644);
        }
    }
    if (log_fd >= 0) {
        char buf[64];
        int n = snprintf(buf, sizeof buf, "%.9f\n", t);
        if (n > 0 && write(log_fd, buf, (size_t)n) < 0) log_fd = -1;
    }
    return t;
}
