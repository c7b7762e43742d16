// dataptr.hpp -- one buffer with a host copy and a device copy, on HIP.
//
// Member API of the reference's DataPtr (fluids5.0/src/dataptr.h:34-71,
// dataptr.cpp).  Device memory is HIP linear memory; the OpenGL flags
// (DT_GLVBO, DT_GLTEX) are accepted and mapped to plain device memory, since
// this path has no renderer.  Failures return FX_E* codes instead of exit().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fluidsx_params.h"

typedef unsigned char uchar;
typedef unsigned short ushort;

// usage types and flags (dataptr.h:17-32)
#define DT_MISC 0
#define DT_UCHAR 1
#define DT_UCHAR3 3
#define DT_UCHAR4 4
#define DT_USHORT 5
#define DT_UINT 6
#define DT_INT 7
#define DT_FLOAT 8
#define DT_FLOAT3 12
#define DT_FLOAT4 16

#define DT_CPU 1
#define DT_CUMEM 2
#define DT_CUARRAY 4
#define DT_GLTEX 8
#define DT_GLVBO 16

class DataPtr {
 public:
  DataPtr() = default;
  // Like the reference, memory is released only by an explicit Clear()
  // (dataptr.cpp:25-28).
  ~DataPtr() = default;

  int Resize(int stride, uint64_t cnt, const char* dat = nullptr, uchar dest_flags = DT_CPU);
  int Append(int stride, uint64_t cnt, const char* dat = nullptr, uchar dest_flags = DT_CPU);
  void SetUsage(uchar dt, uchar flags = DT_MISC, fx_int3 res = fx_int3{-1, -1, -1});
  int ReallocateCPU(uint64_t sz);
  int CopyTo(DataPtr* dest, uchar dest_flags, hipStream_t stream = nullptr);
  int Commit(hipStream_t stream = nullptr);
  int Retrieve(hipStream_t stream = nullptr);
  void Clear();

  static int getStride(uchar dtype);
  static uint64_t getDataSz(uint64_t cnt, int stride) { return cnt * (uint64_t)stride; }

  uint64_t getNum() const { return mNum; }
  uint64_t getMax() const { return mMax; }
  char* getData() const { return mCpu; }
  void* getGPU() const { return mGpu; }
  char* getPtr(uint64_t n) const { return mCpu + n * mStride; }

  uint64_t mNum = 0, mMax = 0, mSize = 0;
  int mStride = 0;
  uchar mRefID = 0, mUseType = DT_MISC, mUseFlags = DT_MISC;
  fx_int3 mUseRes{0, 0, 0};
  char* mCpu = nullptr;
  void* mGpu = nullptr;
  int mGLID = -1;
};
