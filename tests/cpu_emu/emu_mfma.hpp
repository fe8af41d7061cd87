// Wave-wide operations of the MFMA X-engine body (csrc/xmfma_body.hpp) for the
// CPU emulation: every emulated lane is a thread, so an MFMA is emulated by
// depositing each lane's operands in shared arrays, a workgroup barrier, each
// lane computing its own four D values with the lane layout the kernel assumes
//   A: lane l = row l%16, K slice l/16;  B: lane l = column l%16, same slice;
//   D: lane l, register v = D[4*(l/16)+v][l%16]
// and a second barrier before the arrays are reused. FP16 products are exact in
// float; int8 products and sums exact in int32.
#pragma once

#include <cstdint>
#include <cstring>

#include "emu_types.hpp"
#include "forms.hpp"
#include "xmfma_body.hpp"

struct EmuMfmaShared {
    corr::xmfma::Frag a[corr::xmfma::kMaxThreads], b[corr::xmfma::kMaxThreads];
    float xf[corr::xmfma::kMaxThreads];
    int32_t xi[corr::xmfma::kMaxThreads];
};

struct EmuMfmaOps {
    int tid;
    EmuMfmaShared *sh;
    HostSync *sync;

    static float half_at(const corr::xmfma::Frag &f, int e) {
        return corr::forms::f16_bits_to_f32((uint16_t)(f.w[e >> 1] >> (16 * (e & 1))));
    }
    static int32_t i8_at(const corr::xmfma::Frag &f, int e) {
        return (int32_t)(int8_t)(uint8_t)(f.w[e >> 2] >> (8 * (e & 3)));
    }

    corr::xmfma::Acc<float> mma(const corr::xmfma::Frag &a, const corr::xmfma::Frag &b,
                                const corr::xmfma::Acc<float> &c) {
        sh->a[tid] = a;
        sh->b[tid] = b;
        (*sync)();
        const int lane = tid & 63, wb = tid & ~63;
        corr::xmfma::Acc<float> d = c;
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v, col = lane & 15;
            float s = 0.0f;
            for (int g = 0; g < 4; ++g)
                for (int e = 0; e < 8; ++e)
                    s += half_at(sh->a[wb + 16 * g + row], e) * half_at(sh->b[wb + 16 * g + col], e);
            d.v[v] = c.v[v] + s;
        }
        (*sync)();
        return d;
    }

    corr::xmfma::Acc<int32_t> mma(const corr::xmfma::Frag &a, const corr::xmfma::Frag &b,
                                  const corr::xmfma::Acc<int32_t> &c) {
        sh->a[tid] = a;
        sh->b[tid] = b;
        (*sync)();
        const int lane = tid & 63, wb = tid & ~63;
        corr::xmfma::Acc<int32_t> d = c;
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v, col = lane & 15;
            int32_t s = 0;
            for (int g = 0; g < 4; ++g)
                for (int e = 0; e < 16; ++e)
                    s += i8_at(sh->a[wb + 16 * g + row], e) * i8_at(sh->b[wb + 16 * g + col], e);
            d.v[v] = c.v[v] + s;
        }
        (*sync)();
        return d;
    }

    // Four 16x16x4 FP32 products in turn, the i-th on word i of the lanes'
    // fragments (one float of A and of B per lane, K slice l/16 of 4)
    corr::xmfma::Acc<float> mma_f32x4(const corr::xmfma::Frag &a, const corr::xmfma::Frag &b,
                                      const corr::xmfma::Acc<float> &c) {
        sh->a[tid] = a;
        sh->b[tid] = b;
        (*sync)();
        const int lane = tid & 63, wb = tid & ~63;
        corr::xmfma::Acc<float> d = c;
        for (int i = 0; i < 4; ++i)
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v, col = lane & 15;
                float s = 0.0f;
                for (int g = 0; g < 4; ++g) {
                    float x, y;
                    std::memcpy(&x, &sh->a[wb + 16 * g + row].w[i], 4);
                    std::memcpy(&y, &sh->b[wb + 16 * g + col].w[i], 4);
                    s += x * y;
                }
                d.v[v] += s;
            }
        (*sync)();
        return d;
    }

    float xor1(float x) {
        sh->xf[tid] = x;
        (*sync)();
        const float r = sh->xf[tid ^ 1];
        (*sync)();
        return r;
    }

    int32_t xor1(int32_t x) {
        sh->xi[tid] = x;
        (*sync)();
        const int32_t r = sh->xi[tid ^ 1];
        (*sync)();
        return r;
    }
};
