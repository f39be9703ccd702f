// Host-side checks of the C ABI under AddressSanitizer and UndefinedBehavior-
// Sanitizer, with no GPU: argument validation of the constructors, every entry
// point on a NULL handle, the text parsing of sfac_tune_import (malformed,
// overlong and extreme records), sfac_tune_export into short buffers, and the
// serving loop's constructor failing without a device.  Built and run by
// tests/test_host_sanitizers.py; prints "ok" at the end.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sfac.hpp"
#include "sfac_ext.h"

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);       \
            ++fails;                                                      \
        }                                                                 \
    } while (0)

static double zero_fn(double r2, double *dA, void *info) {
    (void)r2;
    (void)info;
    *dA = 0.0;
    return 0.0;
}

int main() {
    double L[6] = {10, 11, 12, 1, -1, 0.5};
    int K[3] = {8, 8, 8};
    // constructors: rejected arguments, no device needed
    CHECK(sfac_ctor(L, K, 3) == nullptr);
    CHECK(std::strstr(sfac_last_error(), "invalid order") != nullptr);
    int Kbad[3] = {8, 0, 8};
    CHECK(sfac_ctor(L, Kbad, 4) == nullptr);
    int Kbig[3] = {8, 8, 5000};
    CHECK(sfac_ctor(L, Kbig, 4) == nullptr);
    (void)sfac_ctor_ex(L, K, 4, 1 << 20);  // unknown flags: an error, never a crash
    // every entry point on a NULL handle reports an error and returns
    double vir[6], out[16];
    double w[2] = {1, -1}, x[6] = {0, 0, 0, 1, 1, 1}, dx[6], phi[2];
    sfac(nullptr, 2, w, x);
    CHECK(get_S(nullptr) == nullptr);
    set_L(nullptr, L);
    CHECK(set_A(nullptr, zero_fn, 1.0, nullptr) != 0);
    CHECK(en(nullptr) == 0.0);
    CHECK(de1(nullptr, vir) == 0.0);
    de2(nullptr, 2, w, x, dx);
    potl(nullptr, 2, x, phi);
    sfac_dtor(nullptr);
    CHECK(sfac_set_A_ewald(nullptr, 0.3, 1.0) != 0);
    CHECK(sfac_table(nullptr, nullptr, nullptr) != 0);
    CHECK(sfac_dev(nullptr, 2, w, x) != 0);
    CHECK(sfac_eval_dev(nullptr, 2, w, x, out, dx, phi) != 0);
    CHECK(sfac_eval_dev_noalloc(nullptr, 2, w, x, out, dx, phi) != 0);
    (void)sfac_eval(nullptr, 2, w, x, vir, dx, phi);
    CHECK(sfac_stream(nullptr) == nullptr);
    CHECK(sfac_sync(nullptr) != 0);
    long long info[12];
    CHECK(sfac_info(nullptr, info) != 0);
    CHECK(sfac_set_transform(nullptr, 0) != 0);
    CHECK(sfac_set_graph(nullptr, 1) != 0);
    CHECK(sfac_batch_create(nullptr, 0) == nullptr);
    void *none[1] = {nullptr};
    CHECK(sfac_batch_create(none, 1) == nullptr);
    CHECK(sfac_batch_set(nullptr, 0, 2, w, x, out, dx, phi) != 0);
    CHECK(sfac_batch_eval(nullptr, nullptr) != 0);
    CHECK(sfac_batch_sync(nullptr) != 0);
    CHECK(sfac_batch_set_graph(nullptr, 1) != 0);
    CHECK(sfac_batch_destroy(nullptr) == 0);  // like free(NULL): nothing to do
    CHECK(sfac_accum_S2(nullptr, out) != 0);
    CHECK(sfac_accum_cross(nullptr, nullptr, out) != 0);
    CHECK(sfac_radial(nullptr, out, 4, 0.1, out, out) != 0);
    CHECK(sfac_server_wait(nullptr, 1, nullptr) != 0);
    CHECK(sfac_server_poll(nullptr, 1) < 0);
    CHECK(sfac_server_stats(nullptr, out) != 0);
    CHECK(sfac_server_reset_stats(nullptr) != 0);
    CHECK(sfac_server_destroy(nullptr) == 0);
    char rep[64];
    CHECK(sfac_autotune(nullptr, 10, nullptr, nullptr, 1, rep, sizeof(rep)) != 0);
    CHECK(sfac_batch_autotune(nullptr, 1, rep, sizeof(rep)) != 0);
    CHECK(sfac_tune_current(nullptr, rep, sizeof(rep)) != 0);
    CHECK(sfac_profile(nullptr, 1) != 0);
    // no device: the serving loop's constructor fails cleanly
    CHECK(sfac_server_create(2, 0.0, 4, 100000) == nullptr);

    // remembered launch shapes: well-formed, malformed, overlong and extreme
    CHECK(sfac_tune_clear() == 0);
    const char *good = "128x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0 1\n"
                       "45x45x45/6/0 5,3,3 - 3,5,3 4x4x0 0x0x0 0\n";
    CHECK(sfac_tune_import(good) == 2);
    std::vector<std::string> bad = {
        "",
        "\n\n\n",
        "x",
        "128x128x128",
        "128x128x128/4",
        "128x128x128/4/3",
        "128x128x128/4/3 16,8",
        "128x128x128/4/3 16,8 16,8 8,16",
        "128x128x128/4/3 16,8 16,8 8,16 8x8x0",
        "128x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0",
        "-1x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0 1",
        "99999999999999999999x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0 1",
        "128x128x128/4/3 2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2,2 - - 8x8x0 8x8x0 1",
        "128x128x128/4/3 16,,8 16,8 8,16 8x8x0 8x8x0 1",
        "128x128x128/4/3 16,8, 16,8 8,16 8x8x0 8x8x0 1",
        "128x128x128/4/3 7,8 16,8 8,16 8x8x0 8x8x0 1",
        "128x128x128/4/3 16,8 16,8 8,16 99999999999x8x0 8x8x0 1",
        "128x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x-5 1",
        "128x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0 7",
        "128x128x128/4/3 16,8 16,8 8,16 8x8x0 8x8x0 1 trailing words here",
        std::string(5000, 'a'),
        std::string("128x128x128/4/3 ") + std::string(3000, '2') + " - - 8x8x0 8x8x0 1",
        "\t128x128x128/4/3\t16,8\t16,8\t8,16\t8x8x0\t8x8x0\t1",
    };
    for (const std::string &t : bad) {
        const int k = sfac_tune_import(t.c_str());
        CHECK(k >= -1 && k <= 1);
    }
    CHECK(sfac_tune_import(nullptr) < 0);
    // export: the size query, then buffers shorter than the text
    const int need = sfac_tune_export(nullptr, 0);
    CHECK(need > 1);
    std::vector<char> buf((size_t)need + 8, 'z');
    for (int len : {1, 2, 7, need / 2, need - 1, need}) {
        std::fill(buf.begin(), buf.end(), 'z');
        (void)sfac_tune_export(buf.data(), len);
        CHECK(buf[(size_t)len] == 'z');  // nothing written past len
        CHECK(std::memchr(buf.data(), 0, (size_t)len) != nullptr);  // terminated
    }
    CHECK(sfac_tune_clear() == 0);
    char one[1];
    CHECK(sfac_tune_export(one, 1) >= 0 && one[0] == 0);
    std::printf(fails ? "failed %d\n" : "ok\n", fails);
    return fails ? 1 : 0;
}
