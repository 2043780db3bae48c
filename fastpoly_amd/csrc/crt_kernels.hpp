// Garner reconstruction kernels (restore_crt, crt_convolution.hpp:88-127):
// the lone pass over three residue arrays and the merged pass over a chunk
// table of records.  Non-template kernels, so only the main translation unit
// (fpxg.hip, through ew.hpp) includes this header.
#pragma once

#include "crt.hpp"

namespace fpxg {

__global__ void k_garner(const uint32_t* __restrict__ c1, const uint32_t* __restrict__ c2,
                         const uint32_t* __restrict__ c3, uint32_t* __restrict__ out, uint32_t n,
                         GarnerConst g) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    garner_one(c1, c2, c3, out, i, g);
}

// Every record of a merged launch through the chunk table (chunk -> record,
// first element), kChunk elements per chunk.
__global__ void __launch_bounds__(256) k_garner_batch(const GarnerRec* __restrict__ recs, const uint2* __restrict__ chunks) {
  const uint2 ch = chunks[blockIdx.x];
  GarnerRec R = recs[ch.x];
  make_uniform(R);
  const uint32_t start = __builtin_amdgcn_readfirstlane(ch.y);
  const uint32_t end = min(R.n, start + kChunk);
  for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) garner_one(R.c1, R.c2, R.c3, R.out, i, R.g);
}

}  // namespace fpxg
