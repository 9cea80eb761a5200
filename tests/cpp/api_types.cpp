// C++ API checks the Python tests cannot reach:
//   * the block-cyclic maps num_loc / indxg2p / indxg2l / indxl2g against a
//     brute-force walk of the distribution (the reference's
//     include/ddla/ddla_desc.h:10-30 semantics);
//   * ddla::vector<T> on the GPU handle (include/ddla/ddla_vector.h:27-110
//     of the reference): growth keeps the elements, shrinking and clear()
//     keep the allocation, moves transfer ownership, reset() releases, bad
//     arguments throw.
// "api_types cpu" runs the index checks only (no GPU); no argument runs both.
// Prints one line per failure and exits non-zero if any.
#include <ddla/ddla.h>
#include <ddla/ddla_vector.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

static int failures = 0;
#define EXPECT(cond)                                                                   \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                \
            ++failures;                                                                \
        }                                                                              \
    } while (0)

static void index_maps()
{
    for (int nb : {1, 2, 3, 7, 64})
        for (int p : {1, 2, 3, 4, 5})
            for (int src = 0; src < p; ++src)
                for (int n : {0, 1, 5, 64, 65, 127, 300}) {
                    // deal the blocks round the processes, starting at src
                    std::vector<int> count(p, 0);
                    int owner = src, left_in_block = nb;
                    for (int g = 0; g < n; ++g) {
                        if (left_in_block == 0) {
                            owner = owner + 1 == p ? 0 : owner + 1;
                            left_in_block = nb;
                        }
                        --left_in_block;
                        const int local = count[owner]++;
                        EXPECT(ddla::indxg2p(g, nb, src, p) == owner);
                        EXPECT(ddla::indxg2l(g, nb, p) == local);
                        EXPECT(ddla::indxl2g(local, nb, owner, src, p) == g);
                    }
                    for (int q = 0; q < p; ++q) EXPECT(ddla::num_loc(n, nb, q, src, p) == count[q]);
                }
}

static void vector_on_gpu()
{
    ddla::DdlaHandle_t h = nullptr;
    DDLA_CHECK(ddla::ddlaInit(h, ddla::DdlaBackend::GPU));
    DDLA_CHECK(ddla::ddlaSet(h, 1, 1, 'R'));
    {
        ddla::vector<double> v(h, 100);
        EXPECT(v.size() == 100 && v.capacity() == 100 && !v.empty() && v.data() != nullptr && v.handle() == h);
        std::vector<double> host(100), back(100, 0.0);
        for (int i = 0; i < 100; ++i) host[i] = 0.5 * i - 7.0;
        DDLA_CHECK(ddla::ddlaMemcpy(h, v.data(), host.data(), 100 * sizeof(double)));
        v.resize(1000);  // growth keeps the first 100 values
        EXPECT(v.size() == 1000 && v.capacity() == 1000);
        DDLA_CHECK(ddla::ddlaMemcpy(h, back.data(), v.data(), 100 * sizeof(double)));
        DDLA_CHECK(ddla::ddlaSynchronize(h));
        EXPECT(std::memcmp(back.data(), host.data(), 100 * sizeof(double)) == 0);
        double* const before = v.data();
        v.resize(10);  // shrinking keeps the allocation
        EXPECT(v.size() == 10 && v.capacity() == 1000 && v.data() == before);
        ddla::vector<double> w(std::move(v));
        EXPECT(w.data() == before && w.size() == 10 && v.data() == nullptr && v.size() == 0 && v.capacity() == 0);
        ddla::vector<double> x(h);
        x = std::move(w);
        EXPECT(x.data() == before && w.data() == nullptr);
        x.clear();
        EXPECT(x.size() == 0 && x.capacity() == 1000 && x.data() == before);
        x.reset();
        EXPECT(x.size() == 0 && x.capacity() == 0 && x.data() == nullptr);
        bool threw = false;
        try {
            x.reserve(x.max_size() + 1);
        } catch (const std::length_error&) {
            threw = true;
        }
        EXPECT(threw);
        threw = false;
        try {
            ddla::vector<float> bad(nullptr);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        EXPECT(threw);
    }
    DDLA_CHECK(ddla::ddlaDestroy(h));
}

int main(int argc, char** argv)
{
    index_maps();
    if (!(argc > 1 && std::strcmp(argv[1], "cpu") == 0)) vector_on_gpu();
    std::printf("%s: %d failure(s)\n", argc > 1 ? argv[1] : "all", failures);
    return failures == 0 ? 0 : 1;
}
