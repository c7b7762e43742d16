// dataptr.cpp -- DataPtr on HIP (behaviour of fluids5.0/src/dataptr.cpp).
#include "dataptr.hpp"

#include <cstdlib>
#include <cstring>

#include "fx_common.hpp"

void DataPtr::Clear() {
  // dataptr.cpp:30-63: free host memory, then device memory.
  if (mCpu != nullptr) {
    free(mCpu);
    mCpu = nullptr;
  }
  if (mGpu != nullptr) {
    (void)hipFree(mGpu);
    mGpu = nullptr;
  }
  mGLID = -1;
  mNum = 0;
  mMax = 0;
  mSize = 0;
}

int DataPtr::getStride(uchar dtype) {
  // dataptr.cpp:65-78
  switch (dtype) {
    case DT_UCHAR: return 1;
    case DT_UCHAR3: return 3;
    case DT_UCHAR4: return 4;
    case DT_USHORT: return 2;
    case DT_UINT: return 4;
    case DT_FLOAT: return 4;
    case DT_FLOAT4: return 16;
    default: return dtype;
  }
}

void DataPtr::SetUsage(uchar dt, uchar flags, fx_int3 res) {
  mUseType = dt;
  if (flags != DT_MISC) mUseFlags = flags;
  if (res.x != -1) mUseRes = res;
}

int DataPtr::ReallocateCPU(uint64_t sz) {
  if (mSize == sz && mCpu != nullptr) return FX_OK;
  char* nd = static_cast<char*>(malloc(sz ? sz : 1));
  if (nd == nullptr) return fx::fail(FX_ERANGE, "DataPtr::ReallocateCPU: malloc(%llu) failed",
                                     (unsigned long long)sz);
  if (mCpu != nullptr) {
    memcpy(nd, mCpu, mSize < sz ? mSize : sz);
    free(mCpu);
  }
  mSize = sz;
  mCpu = nd;
  return FX_OK;
}

int DataPtr::Resize(int stride, uint64_t cnt, const char* dat, uchar dest_flags) {
  Clear();
  return Append(stride, cnt, dat, dest_flags);
}

int DataPtr::Append(int stride, uint64_t cnt, const char* dat, uchar dest_flags) {
  // dataptr.cpp:105-194: grow the host copy and (re)allocate the device copy.
  if (cnt == 0) return FX_OK;
  mStride = stride;
  const uint64_t data_size = getDataSz(cnt, stride);
  const uint64_t old_size = mSize;
  const uint64_t new_size = mSize + data_size;
  if (dest_flags & DT_CPU) {
    FX_TRY(ReallocateCPU(new_size));
    if (dat != nullptr) memcpy(mCpu + old_size, dat, data_size);
  }
  mMax += cnt;
  mSize = new_size;
  if (dest_flags & (DT_CUMEM | DT_GLVBO | DT_GLTEX | DT_CUARRAY)) {
    void* nd = nullptr;
    FX_HIP(hipMalloc(&nd, new_size), "DataPtr::Append");
    if (mGpu != nullptr) {
      if (old_size > 0) FX_HIP(hipMemcpy(nd, mGpu, old_size, hipMemcpyDeviceToDevice), "DataPtr::Append");
      (void)hipFree(mGpu);
    }
    if (dat != nullptr) FX_HIP(hipMemcpy(static_cast<char*>(nd) + old_size, dat, data_size,
                                         hipMemcpyHostToDevice), "DataPtr::Append");
    mGpu = nd;
  }
  return FX_OK;
}

int DataPtr::Commit(hipStream_t stream) {
  // dataptr.cpp:196-225: only the in-use elements go to the device.
  if (mGpu == nullptr || mCpu == nullptr) return FX_OK;
  const uint64_t sz = mNum * (uint64_t)mStride;
  FX_HIP(hipMemcpyAsync(mGpu, mCpu, sz, hipMemcpyHostToDevice, stream), "DataPtr::Commit");
  FX_HIP(hipStreamSynchronize(stream), "DataPtr::Commit");
  return FX_OK;
}

int DataPtr::Retrieve(hipStream_t stream) {
  // dataptr.cpp:228-276: the whole allocation comes back.
  if (mCpu == nullptr) {
    const uint64_t sz = mSize;
    mSize = 0;
    FX_TRY(ReallocateCPU(sz));
  }
  if (mGpu == nullptr) return FX_OK;
  FX_HIP(hipMemcpyAsync(mCpu, mGpu, mSize, hipMemcpyDeviceToHost, stream), "DataPtr::Retrieve");
  FX_HIP(hipStreamSynchronize(stream), "DataPtr::Retrieve");
  return FX_OK;
}

int DataPtr::CopyTo(DataPtr* dest, uchar dest_flags, hipStream_t stream) {
  // dataptr.cpp:278-292; a size mismatch is an error code, not exit(-11).
  if (mSize != dest->mSize) return fx::fail(FX_EMISMATCH, "DataPtr::CopyTo: sizes do not match");
  if ((mUseFlags & DT_CPU) && (dest_flags & DT_CPU) && mCpu && dest->mCpu)
    memcpy(dest->mCpu, mCpu, mSize);
  if (mGpu != nullptr && dest->mGpu != nullptr && (dest_flags & DT_CUMEM))
    FX_HIP(hipMemcpyAsync(dest->mGpu, mGpu, mSize, hipMemcpyDeviceToDevice, stream), "DataPtr::CopyTo");
  return FX_OK;
}
