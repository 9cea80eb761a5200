/* Test double of the three MPI names the C++ header's communicator overloads
 * use (MPI_Comm, MPI_COMM_WORLD, MPI_Comm_size, MPI_SUCCESS), for
 * tests/cpp/api_mpi.cpp only; the image has no MPI.  MPI_Comm is an int, as in
 * MPICH, the harder case for overload resolution against ddlaSet(handle, int,
 * int, char).  The size of MPI_COMM_WORLD is set by the test. */
#ifndef DDLA_TEST_MPI_STUB_H
#define DDLA_TEST_MPI_STUB_H
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#define MPI_COMM_SELF ((MPI_Comm)0x44000001)
#define MPI_SUCCESS 0
#define MPI_ERR_COMM 5
extern int ddla_test_world_size;
inline int MPI_Comm_size(MPI_Comm comm, int* size)
{
    if (comm == MPI_COMM_WORLD) *size = ddla_test_world_size;
    else if (comm == MPI_COMM_SELF) *size = 1;
    else return MPI_ERR_COMM;
    return MPI_SUCCESS;
}
#endif
