// C++ API checks the Python tests cannot reach:
//   * the block-cyclic maps num_loc / indxg2p / indxg2l / indxl2g against a
//     brute-force walk of the distribution (the reference's
//     include/ddla/ddla_desc.h:10-30 semantics);
//   * ddla::vector<T> on the GPU handle (include/ddla/ddla_vector.h:27-110
//     of the reference): growth keeps the elements, shrinking and clear()
//     keep the allocation, moves transfer ownership, reset() releases, bad
//     arguments throw;
//   * the handle helpers of include/ddla/ddla_handle_t.h:146-334 of the
//     reference: pinned host memory, ddlaMemsetTri / ddlaMemcpyTri in every
//     direction (both triangles, with and without the diagonal, padding rows
//     and the other triangle untouched, a 3-byte element type), events, device
//     queries, the rank <-> grid maps and the descriptor-level index maps
//     (include/ddla/ddla_desc.h:143-191).
// "api_types cpu" runs the index checks only (no GPU); no argument runs both.
// Prints one line per failure and exits non-zero if any.
#include <ddla/ddla.h>
#include <ddla/ddla_vector.h>

#include "internal.h"  // white-box: the handle's own look-ahead streams

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

static void handle_helpers_on_gpu(ddla::DdlaHandle_t h);

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
        handle_helpers_on_gpu(h);
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

// Element (i, j) of a column-major matrix with leading dimension ld is in the
// triangle: upper i <= j, lower i >= j, strict without the diagonal.
static bool in_tri(bool upper, bool diag, int i, int j)
{
    return upper ? (diag ? i <= j : i < j) : (diag ? i >= j : i > j);
}

struct Rgb {  // 3-byte element: the byte-by-byte path of the triangular kernels
    unsigned char r, g, b;
};

static void handle_helpers_on_gpu(ddla::DdlaHandle_t h)
{
    using ddla::ddlaStatus_t;
    int count = 0;
    DDLA_CHECK(ddla::ddlaGetDeviceCount(h, &count));
    EXPECT(count >= 1);
    std::size_t free_b = 0, total_b = 0;
    DDLA_CHECK(ddla::ddlaMemGetInfo(h, &free_b, &total_b));
    EXPECT(total_b > 0 && free_b <= total_b);
    DDLA_CHECK(ddla::ddlaSetDevice(h));
    DDLA_CHECK(ddla::ddlaGetLastError(h));
    int row = 9, col = 9;
    ddla::ddlaRankToRc(h, 0, row, col);
    EXPECT(row == 0 && col == 0 && ddla::ddlaRcToRank(h, 0, 0) == 0);
    ddla::ddlaRankToRc(nullptr, 0, row, col);
    EXPECT(row == -1 && col == -1 && ddla::ddlaRcToRank(nullptr, 0, 0) == -1);
    // descriptor-level maps at a 1 x 1 grid: identities inside M x N, -1 beyond
    int desc[ddla::DDLA_DLEN_];
    DDLA_CHECK(ddla::ddlaDescInit(desc, h, 100, 70, 16, 16, 0, 0));
    for (int g = 0; g < 100; ++g) EXPECT(ddla::indx_g2l_r(desc, h, g) == g && ddla::indx_l2g_r(desc, h, g) == g);
    for (int g = 0; g < 70; ++g) EXPECT(ddla::indx_g2l_c(desc, h, g) == g && ddla::indx_l2g_c(desc, h, g) == g);
    EXPECT(ddla::indx_g2l_r(desc, h, 100) == -1 && ddla::indx_g2l_c(desc, h, 70) == -1);
    EXPECT(ddla::indx_g2l_r(desc, nullptr, 0) == -1 && ddla::indx_l2g_c(desc, nullptr, 0) == -1);

    double* pinned = nullptr;
    DDLA_CHECK(ddla::ddlaMallocHost(h, &pinned, 64 * sizeof(double)));
    EXPECT(pinned != nullptr);
    void* none = &count;
    DDLA_CHECK(ddla::ddlaMallocHost(h, &none, 0));
    EXPECT(none == nullptr);
    ddla::ddlaEvent_t ev = nullptr;
    DDLA_CHECK(ddla::ddlaEventCreate(h, &ev));
    auto sync = [&] {
        DDLA_CHECK(ddla::ddlaEventRecord(h, ev));
        DDLA_CHECK(ddla::ddlaEventWait(h, ev));
        DDLA_CHECK(ddla::ddlaEventSynchronize(h, ev));
    };

    // 7 x 5 doubles with leading dimension 9 (two padding rows per column)
    const int m = 7, n = 5, ld = 9;
    const std::size_t bytes = (std::size_t)ld * n * sizeof(double);
    std::vector<double> host(ld * n), back(ld * n), sentinel(ld * n, -1.0);
    for (int i = 0; i < ld * n; ++i) host[i] = 100.0 + i;
    ddla::vector<double> dev(h, ld * n), dev2(h, ld * n);
    for (int strict = 0; strict < 2; ++strict)
        for (char uplo : {'U', 'l'}) {
            const bool upper = uplo == 'U', diag = strict == 0;
            // memset: zero the triangle of a copy of host
            DDLA_CHECK(ddla::ddlaMemcpy(h, dev.data(), host.data(), bytes));
            DDLA_CHECK(ddla::ddlaMemsetTri(h, uplo, m, n, dev.data(), ld, 0, diag));
            DDLA_CHECK(ddla::ddlaMemcpy(h, back.data(), dev.data(), bytes));
            sync();
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < ld; ++i)
                    EXPECT(back[i + j * ld] == ((i < m && in_tri(upper, diag, i, j)) ? 0.0 : host[i + j * ld]));
            // host -> device, device -> device, device -> host, host -> host copies of the triangle
            DDLA_CHECK(ddla::ddlaMemcpy(h, dev2.data(), sentinel.data(), bytes));
            DDLA_CHECK(ddla::ddlaMemcpyTri(h, uplo, m, n, dev2.data(), ld, host.data(), ld, diag));
            DDLA_CHECK(ddla::ddlaMemcpy(h, dev.data(), sentinel.data(), bytes));
            DDLA_CHECK(ddla::ddlaMemcpyTri(h, uplo, m, n, dev.data(), ld, dev2.data(), ld, diag));
            std::vector<double> out(ld * n, -2.0), hh(ld * n, -3.0);
            DDLA_CHECK(ddla::ddlaMemcpyTri(h, uplo, m, n, out.data(), ld, dev.data(), ld, diag));
            sync();
            DDLA_CHECK(ddla::ddlaMemcpyTri(h, uplo, m, n, hh.data(), ld, out.data(), ld, diag));
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < ld; ++i) {
                    const bool sel = i < m && in_tri(upper, diag, i, j);
                    EXPECT(out[i + j * ld] == (sel ? host[i + j * ld] : -2.0));
                    EXPECT(hh[i + j * ld] == (sel ? host[i + j * ld] : -3.0));
                }
        }
    // memset repeats the low byte: value 0x101 on floats writes 0x01010101
    {
        const int fm = 33, fn = 4;
        std::vector<float> f(fm * fn, 2.5f), fb(fm * fn);
        ddla::vector<float> fd(h, fm * fn);
        DDLA_CHECK(ddla::ddlaMemcpy(h, fd.data(), f.data(), f.size() * sizeof(float)));
        DDLA_CHECK(ddla::ddlaMemsetTri(h, 'L', fm, fn, fd.data(), fm, 0x101));
        DDLA_CHECK(ddla::ddlaMemcpy(h, fb.data(), fd.data(), f.size() * sizeof(float)));
        sync();
        unsigned int pattern = 0x01010101u;
        float want;
        std::memcpy(&want, &pattern, sizeof(want));
        for (int j = 0; j < fn; ++j)
            for (int i = 0; i < fm; ++i)
                EXPECT(std::memcmp(&fb[i + j * fm], i >= j ? &want : &f[i + j * fm], sizeof(float)) == 0);
    }
    // 3-byte elements (byte-by-byte kernel path), device to device
    {
        const int rm = 6, rn = 6, rl = 6;
        std::vector<Rgb> src(rl * rn), dst(rl * rn, Rgb{1, 2, 3}), got(rl * rn);
        for (int i = 0; i < rl * rn; ++i) src[i] = Rgb{(unsigned char)i, (unsigned char)(i + 1), (unsigned char)(i + 2)};
        ddla::vector<Rgb> rs(h, rl * rn), rd(h, rl * rn);
        DDLA_CHECK(ddla::ddlaMemcpy(h, rs.data(), src.data(), src.size() * sizeof(Rgb)));
        DDLA_CHECK(ddla::ddlaMemcpy(h, rd.data(), dst.data(), dst.size() * sizeof(Rgb)));
        DDLA_CHECK(ddla::ddlaMemcpyTri(h, 'U', rm, rn, rd.data(), rl, rs.data(), rl, false));
        DDLA_CHECK(ddla::ddlaMemcpy(h, got.data(), rd.data(), got.size() * sizeof(Rgb)));
        sync();
        for (int j = 0; j < rn; ++j)
            for (int i = 0; i < rm; ++i) {
                const Rgb& e = got[i + j * rl];
                const Rgb& w = i < j ? src[i + j * rl] : dst[i + j * rl];
                EXPECT(e.r == w.r && e.g == w.g && e.b == w.b);
            }
    }
    // argument rules
    EXPECT(ddla::ddlaMemsetTri(h, 'X', m, n, dev.data(), ld) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaMemsetTri(h, 'U', m, n, dev.data(), m - 1) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaMemsetTri(h, 'U', -1, n, dev.data(), ld) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaMemsetTri(h, 'U', m, n, host.data(), ld) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaMemsetTri(h, 'U', 0, n, static_cast<double*>(nullptr), 0) == ddlaStatus_t::DDLA_STATUS_SUCCESS);
    EXPECT(ddla::ddlaMemsetTri(h, 'U', m, 1, static_cast<double*>(nullptr), 0, 0, false) ==
           ddlaStatus_t::DDLA_STATUS_SUCCESS);
    EXPECT(ddla::ddlaMemcpyTri(h, 'L', m, n, dev.data(), ld, static_cast<const double*>(nullptr), ld) ==
           ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaMemcpyTri(nullptr, 'L', m, n, dev.data(), ld, dev2.data(), ld) ==
           ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE);
    DDLA_CHECK(ddla::ddlaFreeHost(h, pinned));
    DDLA_CHECK(ddla::ddlaFreeHost(h, nullptr));
    DDLA_CHECK(ddla::ddlaEventDestroy(h, ev));
    DDLA_CHECK(ddla::ddlaDeviceSynchronize(h));
    // ddlaSetStream refuses the handle's private look-ahead streams (the
    // reference returns INVALID_VALUE for a queue the handle owns) and keeps
    // its compute stream.
    void* before = ddla::ddlaGetStream(h);
    EXPECT(ddla::ddlaSetStream(h, static_cast<void*>(h->side)) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaSetStream(h, static_cast<void*>(h->look)) == ddlaStatus_t::DDLA_STATUS_INVALID_VALUE);
    EXPECT(ddla::ddlaGetStream(h) == before);
}

int main(int argc, char** argv)
{
    index_maps();
    if (!(argc > 1 && std::strcmp(argv[1], "cpu") == 0)) vector_on_gpu();
    std::printf("%s: %d failure(s)\n", argc > 1 ? argv[1] : "all", failures);
    return failures == 0 ? 0 : 1;
}
