// Host sanitizer driver: runs the descriptor validator (umi_amd/csrc/validate.hip,
// host code) and every function of the CPU oracle (oracle/umi_oracle.c) on
// small, ragged inputs held in heap buffers of exactly the size each call
// needs, so that AddressSanitizer reports any read or write outside them and
// UndefinedBehaviorSanitizer any overflow or bad shift.  Built and run by
// tests/test_host_sanitizers.py with g++/gcc (no GPU).
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "umi_hip.h"

namespace umi {
umi_status set_error(umi_status code, const char *fmt, ...) {  // the library's error sink, host only
    va_list a;
    va_start(a, fmt);
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, a);
    va_end(a);
    return code;
}
}  // namespace umi

extern "C" {
void oracle_set_threads(int n);
void oracle_me_full_search_u8(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int32_t *, int32_t *);
void oracle_sad_volume_i32(const int32_t *, const int32_t *, int, int, int, int, int, int32_t *);
void oracle_cost_volume_u8(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int, int, int, int32_t *);
void oracle_stereo_wta_u8(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int, int, int, int32_t *);
void oracle_box_sum_u8(const uint8_t *, int, int, int, int, int, uint32_t *);
void oracle_haar_u8(const uint8_t *, int, int, int, const int32_t *, int, int, int, int32_t *);
void oracle_me_subpel_u8(const uint8_t *, const uint8_t *, int, int, int, const int32_t *, int32_t *);
void oracle_filter3x3_i32(const int32_t *, int, int, int, int32_t, int32_t, int32_t, int, int32_t *, int);
void oracle_filter3x3_f32(const float *, int, int, int, float, float, float, int, float *, int);
void oracle_wfilter3x3_i32(const int32_t *, int, int, int, const int32_t *, int32_t, int32_t, int, int32_t *, int);
void oracle_wfilter3x3_f32(const float *, int, int, int, const float *, float, float, int, float *, int);
void oracle_wfilter3x3_f64(const float *, int, int, int, const float *, float, float, int, double *, double *);
void oracle_bilateral_f32(const float *, int, int, int, int, const float *, float, float, int, float *, int);
void oracle_bilateral_f64(const float *, int, int, int, int, const float *, float, float, int, double *);
void oracle_iimg_u8_u32(const uint8_t *, int, int, int, uint32_t *);
void oracle_iimg_i32(const int32_t *, int, int, int, int32_t *);
void oracle_iimg_f64(const float *, int, int, int, double *);
void oracle_gemm_f32_rows(const float *, const float *, float *, int, int, int, int, int);
void oracle_conv_f32(const float *, const float *, float *, int, int, int, int, int, int, int, int, int, int, int);
int oracle_nv12_chroma_pred(const uint8_t *, int, int, int, int, int, int, int, int);
void oracle_me_nv12_u8(const uint8_t *, const uint8_t *, const uint8_t *, const uint8_t *, int, int, int, int, int,
                       int32_t *);
void oracle_sgm_u8(const uint8_t *, const uint8_t *, int, int, int, int, int, int, int, int, int, int, int32_t *);
}

static std::mt19937 g(7);
template <class T>
static std::vector<T> rnd(size_t n, int lo, int hi) {
    std::uniform_int_distribution<int> d(lo, hi);
    std::vector<T> v(n);
    for (auto &x : v) x = (T)d(g);
    return v;
}

static int validator_cases() {
    int refused = 0;
    for (int n_arr = 0; n_arr <= 8; ++n_arr) {
        for (int variant = 0; variant < 6; ++variant) {
            int32_t p[16] = {1, 1, 1, 1, 1, 1, 1, 1, 300, 16, 4, 1, 500, 16, 8, 1}, a[8] = {1, 1, 1, 1, 3, 3, 3, 3};
            if (variant == 1) p[10] = 0;            // warp 0
            if (variant == 2) p[11] = 3;            // shuffle not a power of two
            if (variant == 3) p[9] = 2048;          // block too large
            if (variant == 4) a[6] = -1;            // negative total
            std::vector<int32_t> views(24 * n_arr + (n_arr == 0), 0), phys(8 * n_arr + (n_arr == 0), 1),
                planes(n_arr + (n_arr == 0), 1);
            for (int i = 0; i < n_arr; ++i) phys[8 * i + 1] = phys[8 * i + 3] = phys[8 * i + 5] = phys[8 * i + 7] = 0;
            if (variant == 5 && n_arr) views[3] = 9;  // view dim out of range
            const int n_in = n_arr / 2, n_out = n_arr - n_in;
            refused += umi_validate_execute(p, a, n_in, n_out, views.data(), phys.data(), planes.data(), 0x1122,
                                            n_arr ? 0xc0 : 0) != UMI_OK;
        }
    }
    return refused;
}

int main() {
    std::printf("validator: %d refusals\n", validator_cases());
    oracle_set_threads(1);
    {  // ME with a ragged frame and a window wider than the frame
        const int H = 13, W = 21, B = 8, R = 9, S = 2 * R + 1, nby = H / B, nbx = W / B;
        auto c = rnd<uint8_t>(H * W, 0, 255), r = rnd<uint8_t>(H * W, 0, 255);
        std::vector<int32_t> mv(nby * nbx * 3), vol((size_t)nby * nbx * S * S);
        oracle_me_full_search_u8(c.data(), r.data(), H, W, W, B, R, 0, nby, mv.data(), vol.data());
        auto ci = rnd<int32_t>(H * W, -1000, 1000);
        std::vector<int32_t> v2((size_t)(H / 4) * (W / 4) * 5 * 5);
        oracle_sad_volume_i32(ci.data(), ci.data(), H, W, 4, 5, -2, v2.data());
    }
    {  // cost volume, winner-take-all, SGM with shifts past the image
        const int H = 5, W = 9, D = 12;
        auto L = rnd<uint8_t>(H * W, 0, 255), R = rnd<uint8_t>(H * W, 0, 255);
        std::vector<int32_t> cv((size_t)H * W * D), wta((size_t)H * W * 2), sgm((size_t)H * W * 2);
        oracle_cost_volume_u8(L.data(), R.data(), H, W, W, D, -3, 1, 3, 3, 0, H, cv.data());
        oracle_stereo_wta_u8(L.data(), R.data(), H, W, W, D, 20, -1, 3, 0, 0, H, wta.data());
        oracle_sgm_u8(L.data(), R.data(), H, W, W, D, 5, 1, 3, 3, 10, 120, sgm.data());
        oracle_sgm_u8(L.data(), R.data(), 1, 1, 1, 1, 0, -1, 0, 0, 0, 0, sgm.data());
    }
    {  // box sums, rectangle features, integral images
        const int H = 7, W = 11;
        auto in = rnd<uint8_t>(H * W, 0, 255);
        std::vector<uint32_t> box(H * W), ii(H * W);
        oracle_box_sum_u8(in.data(), H, W, W, 20, 1, box.data());
        oracle_iimg_u8_u32(in.data(), H, W, W, ii.data());
        const int32_t feats[2 * 15] = {0, 0, 5, 3, 1, 0, 3, 5, 3, -2, 1, 1, 1, 1, 7,
                                       2, 2, 3, 3, -1, 0, 0, 1, 1, 0, 4, 5, 1, 1, 3};
        const int wh = 6, ww = 5;
        std::vector<int32_t> hf(2 * (H - wh + 1) * (W - ww + 1));
        oracle_haar_u8(in.data(), H, W, W, feats, 2, wh, ww, hf.data());
        auto i32 = rnd<int32_t>(H * W, -100000, 100000);
        std::vector<int32_t> o32(H * W);
        oracle_iimg_i32(i32.data(), H, W, W, o32.data());
        std::vector<float> f(1, 3.5f);
        std::vector<double> o64(1);
        oracle_iimg_f64(f.data(), 1, 1, 1, o64.data());
    }
    {  // filters and bilateral, both boundary modes, pitched input
        const int H = 4, W = 5, ip = 6;
        auto i = rnd<int32_t>(H * ip, -50, 50);
        std::vector<float> f(H * ip);
        for (int k = 0; k < H * ip; ++k) f[k] = (float)i[k] * 0.25f;
        const int32_t wi[9] = {1, 2, 1, 2, 4, 2, 1, 2, 1};
        const float wf[9] = {0.1f, 0.2f, 0.1f, 0.2f, 0.4f, 0.2f, 0.1f, 0.2f, 0.1f};
        std::vector<int32_t> oi(H * W);
        std::vector<float> of(H * W), ws(7 * 7, 0.5f);
        std::vector<double> od(H * W), mag(H * W);
        for (int mode = 0; mode < 2; ++mode) {
            oracle_filter3x3_i32(i.data(), H, W, ip, -5, 2, 999, mode, oi.data(), W);
            oracle_filter3x3_f32(f.data(), H, W, ip, -5.f, 2.f, 9.f, mode, of.data(), W);
            oracle_wfilter3x3_i32(i.data(), H, W, ip, wi, -5, 56, mode, oi.data(), W);
            oracle_wfilter3x3_f32(f.data(), H, W, ip, wf, -5.f, 5.f, mode, of.data(), W);
            oracle_wfilter3x3_f64(f.data(), H, W, ip, wf, -5.f, 5.f, mode, od.data(), mag.data());
            oracle_bilateral_f32(f.data(), H, W, ip, 3, ws.data(), -0.5f, 1.f, mode, of.data(), W);
            oracle_bilateral_f64(f.data(), H, W, ip, 3, ws.data(), -0.5f, 1.f, mode, od.data());
        }
    }
    {  // quarter-pel refinement with vectors at the frame edge
        const int H = 16, W = 24, nb = (H / 8) * (W / 8);
        auto c = rnd<uint8_t>(H * W, 0, 255), r = rnd<uint8_t>(H * W, 0, 255);
        std::vector<int32_t> mv(nb * 3), out(nb * 3);
        for (int b = 0; b < nb; ++b) mv[3 * b] = (b & 1) ? -20 : 20, mv[3 * b + 1] = (b & 2) ? 17 : -17;
        oracle_me_subpel_u8(c.data(), r.data(), H, W, W, mv.data(), out.data());
    }
    {  // GEMM rows and a strided, padded convolution
        const int M = 3, N = 5, K = 7;
        std::vector<float> A(M * K, 1.f), Bm(K * N, 2.f), Cm(M * N);
        oracle_gemm_f32_rows(A.data(), Bm.data(), Cm.data(), M, N, K, 1, 3);
        const int n = 2, ch = 3, h = 5, w = 6, k = 4, r = 3, s = 2, stride = 2, pad = 1;
        const int oh = (h + 2 * pad - r) / stride + 1, ow = (w + 2 * pad - s) / stride + 1;
        std::vector<float> x(n * ch * h * w, 0.5f), wt(k * ch * r * s, 0.25f), o(n * k * oh * ow);
        oracle_conv_f32(x.data(), wt.data(), o.data(), n, ch, h, w, k, r, s, stride, pad, 0, n);
    }
    {  // NV12: the predictor far outside the plane, and the search on an 8x10 frame
        const int H = 8, W = 10;
        auto cy = rnd<uint8_t>(H * W, 0, 255), ry = rnd<uint8_t>(H * W, 0, 255);
        auto cuv = rnd<uint8_t>(H / 2 * W, 0, 255), ruv = rnd<uint8_t>(H / 2 * W, 0, 255);
        int acc = 0;
        for (int Y = -40; Y <= 40; Y += 7)
            for (int X = -40; X <= 40; X += 9) acc += oracle_nv12_chroma_pred(ruv.data(), H / 2, W / 2, W, 1, Y, X, 4, 4);
        std::vector<int32_t> out((H / 8) * (W / 8) * 3);
        oracle_me_nv12_u8(cy.data(), cuv.data(), ry.data(), ruv.data(), H, W, W, W, 16, out.data());
        std::printf("nv12 predictor sum %d\n", acc);
    }
    std::printf("sanitized run complete\n");
    return 0;
}
