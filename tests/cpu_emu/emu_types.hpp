// Host stand-ins for the HIP vector types the kernel bodies use, so the very
// same kernel-body headers compile for a CPU emulation of one workgroup.
#pragma once
#include <barrier>

struct float2 {
    float x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
struct char2 {
    signed char x, y;
};
struct int2 {
    int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }
struct float4 {
    float x, y, z, w;
};
struct uint2 {
    unsigned x, y;
};
struct uint4 {
    unsigned x, y, z, w;
};
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }

#ifndef CORR_HD
#define CORR_HD
#endif

// The workgroup barrier of the emulation: every emulated lane is a thread.
struct HostSync {
    std::barrier<> *b;
    void operator()() { b->arrive_and_wait(); }
};
