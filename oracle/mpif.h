c     Serial stand-in for mpif.h used only by the oracle build of the
c     reference (test infrastructure); values pair with mpi_serial.c.
      integer MPI_COMM_WORLD, MPI_STATUS_SIZE
      integer MPI_INTEGER, MPI_REAL, MPI_REAL8, MPI_LOGICAL
      integer MPI_ANY_TAG, MPI_ANY_SOURCE, MPI_DOUBLE_PRECISION
      integer MPI_CHARACTER
      parameter (MPI_COMM_WORLD=91, MPI_STATUS_SIZE=6)
      parameter (MPI_INTEGER=1, MPI_REAL=2, MPI_REAL8=3)
      parameter (MPI_LOGICAL=4, MPI_ANY_TAG=-1, MPI_ANY_SOURCE=-2)
      parameter (MPI_DOUBLE_PRECISION=3, MPI_CHARACTER=5)
      double precision MPI_WTIME
      external MPI_WTIME
