/*
 * Serial (one-rank) stand-in for the handful of MPI routines the reference
 * miniSMAC2D Fortran calls.  TEST INFRASTRUCTURE ONLY: it is linked into the
 * reference oracle build under oracle/_ref/ so that the unmodified reference
 * sources run as a single rank on a machine without an MPI installation.
 *
 * With one rank every point-to-point call in the reference is unreachable
 * (each is guarded by `numprocs.gt.1` or a node-id test); they abort loudly
 * if ever reached.  Broadcasts from rank 0 to rank 0 are no-ops.
 *
 * Fortran binding: lower case with one trailing underscore (flang default),
 * every argument by reference.  Constants match oracle/mpif.h.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static void unreachable(const char *what)
{
    fprintf(stderr, "mpi_serial: %s reached with one rank\n", what);
    abort();
}

void mpi_init_(int *ierr) { *ierr = 0; }
void mpi_finalize_(int *ierr) { *ierr = 0; }
void mpi_comm_rank_(int *comm, int *rank, int *ierr) { (void)comm; *rank = 0; *ierr = 0; }
void mpi_comm_size_(int *comm, int *size, int *ierr) { (void)comm; *size = 1; *ierr = 0; }
void mpi_barrier_(int *comm, int *ierr) { (void)comm; *ierr = 0; }

void mpi_bcast_(void *buf, int *count, int *type, int *root, int *comm, int *ierr)
{
    (void)buf; (void)count; (void)type; (void)comm;
    if (*root != 0) unreachable("MPI_BCAST from a non-zero root");
    *ierr = 0;
}

void mpi_send_(void *buf, int *count, int *type, int *dest, int *tag, int *comm, int *ierr)
{
    (void)buf; (void)count; (void)type; (void)dest; (void)tag; (void)comm; (void)ierr;
    unreachable("MPI_SEND");
}

void mpi_recv_(void *buf, int *count, int *type, int *src, int *tag, int *comm, int *status,
               int *ierr)
{
    (void)buf; (void)count; (void)type; (void)src; (void)tag; (void)comm; (void)status;
    (void)ierr;
    unreachable("MPI_RECV");
}

void mpi_probe_(int *src, int *tag, int *comm, int *status, int *ierr)
{
    (void)src; (void)tag; (void)comm; (void)status; (void)ierr;
    unreachable("MPI_PROBE");
}

void mpi_get_count_(int *status, int *type, int *count, int *ierr)
{
    (void)status; (void)type; (void)count; (void)ierr;
    unreachable("MPI_GET_COUNT");
}

void mpi_abort_(int *comm, int *code, int *ierr)
{
    (void)comm; (void)ierr;
    fprintf(stderr, "mpi_serial: MPI_ABORT(%d)\n", *code);
    exit(*code ? *code : 1);
}

double mpi_wtime_(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
