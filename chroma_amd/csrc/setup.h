// Set-up kernels (layout conversion, gauge packing, clover construction); implemented in setup.hip.
#pragma once
#include "cmx_internal.h"

namespace cmx {

// host order [mu][full site][18] (already on the device) -> FP64 SoA per parity
void gauge_pack64(const Geom& g, const double* dev_host_order, double2* g0, double2* g1, hipStream_t s);
// FP64 SoA (one parity, 4 directions) -> sloppy storage
void gauge_pack_sloppy(const Geom& g, int prec, int recon, const double2* src, void* dst, hipStream_t s);
// mesField + makeClov on all sites in FP64; Fscratch holds 6*9*V double2
void clover_build(const Geom& g, const double2* g0, const double2* g1, double2* Fscratch, const double coef[6],
                  double diag_mass, double2* c0, double2* c1, hipStream_t s);
// LDL^dag inverse of one parity (FP64); per-site log|det| and negative-pivot indicator
void clover_invert(const Geom& g, const double2* src, double2* dst, double* logdet, double* nneg, hipStream_t s);
void clover_convert(int Vh, int prec, const double2* src, void* dst, hipStream_t s);
void clover_unpack(int Vh, const double2* src, double* out, hipStream_t s);
void spinor_pack(int prec, int n, const double* host_dev, void* v, float* nrm, int stride, hipStream_t s);
// hash noise in [-1/2, 1/2) per component (the launch-shape tuner's input: timings depend on the data)
void spinor_noise(int prec, int n, unsigned seed, void* v, float* nrm, int stride, hipStream_t s);
void spinor_unpack(int prec, int n, const void* v, const float* nrm, int stride, double* host_dev, hipStream_t s);
void stream_copy(long n_float4, const void* a, void* b, hipStream_t s);
// streaming reference with the hopping term's read/write mix: out[i] = a[0][i] ^ ... ^ a[4][i] over
// n 16-byte elements (five read streams, one write stream, 5/6 of the bytes read)
void stream_mix(long n_uint4, const void* const a[5], void* out, hipStream_t s);
int sum_partials(int n, const double* a, double* partials, hipStream_t s);

}  // namespace cmx
