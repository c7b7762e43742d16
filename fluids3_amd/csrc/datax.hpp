// datax.hpp -- a set of DataPtr buffers indexed by user id, mirrored into a
// 16-pointer device table.  Member API of the reference's DataX
// (fluids5.0/src/datax.h:66-200) for the buffer, transfer and GPU-access
// parts; the heap allocator (datax.cpp:271-476) is not on the step's path and
// is not provided.
#pragma once

#include <string>
#include <vector>

#include "dataptr.hpp"

#define REF_MAX 64
#define BUNDEF 65535

class DataX {
 public:
  DataX();
  ~DataX();

  bool isActive(int buf) const { return mRef[buf] != BUNDEF; }
  int AddBuffer(int ref, const std::string& name, ushort stride, uint64_t maxcnt,
                uchar dest_flags = DT_CPU);
  void SetBufferUsage(int i, uchar dt, uchar flags = DT_MISC, fx_int3 res = fx_int3{-1, -1, -1});
  void SetNum(uint64_t num);
  void DeleteAllBuffers();
  int CopyAllBuffers(DataX* dest, uchar dest_flags = DT_CPU, hipStream_t stream = nullptr);
  int MatchAllBuffers(DataX* src, uchar use_flags = DT_MISC);
  // Exchanges the device allocations of all buffers with another DataX of
  // identical layout (O(1) replacement of a full device-to-device copy).
  int SwapGPU(DataX* other);

  // GPU access (datax.cpp:68-94)
  int AssignToGPU(const std::string& var_name, hipModule_t module);
  int UpdateGPUAccess();
  void FillTable(cuDataX* table) const;
  void* SymbolAddress() const { return cuData; }

  int Retrieve(int i, hipStream_t stream = nullptr);
  int Commit(int i, hipStream_t stream = nullptr);
  int CommitAll(hipStream_t stream = nullptr);

  int GetNumBuf() const { return (int)mBuf.size(); }
  DataPtr* GetBuffer(int i) { int b = mRef[i]; return b == BUNDEF ? nullptr : &mBuf[b]; }
  uint64_t GetNumElem(int i) const { int b = mRef[i]; return b == BUNDEF ? 0 : mBuf[b].mNum; }
  uint64_t GetMaxElem(int i) const { int b = mRef[i]; return b == BUNDEF ? 0 : mBuf[b].mMax; }
  int GetBufStride(int i) const { int b = mRef[i]; return b == BUNDEF ? 0 : mBuf[b].mStride; }

  // host-side accessors (datax.h:147-162)
  char* bufC(int i, uint64_t n = 0) { int b = mRef[i]; return b == BUNDEF ? nullptr : mBuf[b].mCpu + n; }
  int* bufI(int i, uint64_t n = 0) { int b = mRef[i]; return b == BUNDEF ? nullptr : (int*)mBuf[b].mCpu + n; }
  unsigned* bufUI(int i, uint64_t n = 0) { int b = mRef[i]; return b == BUNDEF ? nullptr : (unsigned*)mBuf[b].mCpu + n; }
  float* bufF(int i, uint64_t n = 0) { int b = mRef[i]; return b == BUNDEF ? nullptr : (float*)mBuf[b].mCpu + n; }
  fx_float3* bufF3(int i, uint64_t n = 0) { int b = mRef[i]; return b == BUNDEF ? nullptr : (fx_float3*)mBuf[b].mCpu + n; }
  void* gpu(int i) const { int b = mRef[i]; return b == BUNDEF ? nullptr : mBuf[b].mGpu; }
  char* cpu(int i) const { int b = mRef[i]; return b == BUNDEF ? nullptr : mBuf[b].mCpu; }

  std::vector<DataPtr> mBuf;
  int mRef[REF_MAX];
  std::string cuDataName;
  void* cuData = nullptr;  // device address of the cuDataX symbol
};
