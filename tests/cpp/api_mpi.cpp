// The communicator overloads of ddlaSet (the reference's
// include/ddla/ddla_handle_t.h:58-64), compiled the way an MPI caller compiles
// against include/ddla/ddla.h (DDLA_USE_MPI, with the test double of
// tests/cpp/mpi_stub/mpi.h standing in for the MPI headers):
//   * a communicator of more than one rank is refused with
//     std::invalid_argument before any device work (null handle, no GPU);
//   * a one-rank communicator forwards to the 1 x 1 grid (a null handle then
//     reports DDLA_STATUS_INVALID_HANDLE);
//   * an invalid communicator gives DDLA_STATUS_INVALID_VALUE;
//   * integer grid shapes still reach the grid-shape overload.
// "api_mpi gpu" also sets up real handles through ddlaSet(h), ddlaSet(h, comm,
// major) and ddlaSet(h, comm, 1, 1) and runs ddlaSynchronize on each.
#include <ddla/ddla.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>

int ddla_test_world_size = 1;

static int failures = 0;
#define EXPECT(cond)                                                                   \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                \
            ++failures;                                                                \
        }                                                                              \
    } while (0)

template <class F>
static bool throws_invalid(F f)
{
    try {
        f();
    } catch (const std::invalid_argument&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

int main(int argc, char** argv)
{
    using ddla::ddlaStatus_t;
    ddla::DdlaHandle_t none = nullptr;
    ddla_test_world_size = 2;
    EXPECT(throws_invalid([&] { ddla::ddlaSet(none); }));
    EXPECT(throws_invalid([&] { ddla::ddlaSet(none, MPI_COMM_WORLD, 'C'); }));
    EXPECT(throws_invalid([&] { ddla::ddlaSet(none, MPI_COMM_WORLD, 2, 1); }));
    EXPECT(ddla::ddlaSet(none, MPI_COMM_SELF) == ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE);
    EXPECT(ddla::ddlaSet(none, (MPI_Comm)7) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    ddla_test_world_size = 1;
    EXPECT(ddla::ddlaSet(none) == ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE);
    EXPECT(ddla::ddlaSet(none, MPI_COMM_WORLD, 1, 1, 'R') == ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE);
    EXPECT(ddla::ddlaSet(none, 1, 1) == ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE);  // grid-shape overload
    if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) {
        for (int form = 0; form < 3; ++form) {
            ddla::DdlaHandle_t h = nullptr;
            DDLA_CHECK(ddla::ddlaInit(h, ddla::DdlaBackend::GPU));
            ddlaStatus_t st = form == 0   ? ddla::ddlaSet(h)
                              : form == 1 ? ddla::ddlaSet(h, MPI_COMM_WORLD, 'C')
                                          : ddla::ddlaSet(h, MPI_COMM_WORLD, 1, 1);
            EXPECT(st == ddlaStatus_t::DDLA_STATUS_SUCCESS);
            EXPECT(ddla::ddlaGetSize(h) == 1 && ddla::ddlaGetRank(h) == 0);
            EXPECT(ddla::ddlaSynchronize(h) == ddlaStatus_t::DDLA_STATUS_SUCCESS);
            DDLA_CHECK(ddla::ddlaDestroy(h));
        }
    }
    std::printf("%s: %d failure(s)\n", argc > 1 ? argv[1] : "cpu", failures);
    return failures == 0 ? 0 : 1;
}
