// datax.cpp -- DataX on HIP (behaviour of fluids5.0/src/datax.cpp:9-138).
#include "datax.hpp"

#include <cstring>

#include "fx_common.hpp"

DataX::DataX() {
  for (int n = 0; n < REF_MAX; n++) mRef[n] = BUNDEF;
}

DataX::~DataX() { DeleteAllBuffers(); }

void DataX::DeleteAllBuffers() {
  for (auto& b : mBuf) b.Clear();
  mBuf.clear();
  for (int n = 0; n < REF_MAX; n++) mRef[n] = BUNDEF;
}

int DataX::AddBuffer(int userid, const std::string& name, ushort stride, uint64_t maxcnt,
                     uchar dest_flags) {
  // datax.cpp:43-55
  (void)name;
  if (userid < 0 || userid >= REF_MAX) return fx::fail(FX_EINVAL, "DataX::AddBuffer: bad id %d", userid);
  DataPtr buf;
  buf.mRefID = (uchar)userid;
  buf.SetUsage(DT_MISC, dest_flags, fx_int3{(int)maxcnt, 1, 1});
  int st = buf.Resize(stride, maxcnt, nullptr, dest_flags);
  if (st != FX_OK) {
    buf.Clear();
    return st;
  }
  mBuf.push_back(buf);
  mRef[userid] = (int)mBuf.size() - 1;
  return FX_OK;
}

void DataX::SetBufferUsage(int i, uchar dt, uchar flags, fx_int3 res) {
  int b = mRef[i];
  if (b == BUNDEF) return;
  mBuf[b].SetUsage(dt, flags, res);
}

void DataX::SetNum(uint64_t num) {
  for (auto& b : mBuf) b.mNum = num;
}

int DataX::AssignToGPU(const std::string& var_name, hipModule_t module) {
  // datax.cpp:68-80; the size check returns FX_EMISMATCH instead of exit(-12).
  size_t len = 0;
  hipDeviceptr_t p = nullptr;
  FX_HIP(hipModuleGetGlobal(&p, &len, module, var_name.c_str()), "DataX::AssignToGPU");
  if (len != sizeof(cuDataX))
    return fx::fail(FX_EMISMATCH, "DataX::AssignToGPU: symbol %s has %zu bytes, cuDataX has %zu",
                    var_name.c_str(), len, sizeof(cuDataX));
  cuDataName = var_name;
  cuData = p;
  return FX_OK;
}

void DataX::FillTable(cuDataX* t) const {
  memset(t, 0, sizeof(*t));
  for (const auto& b : mBuf) t->mbuf[b.mRefID] = (uint64_t)(uintptr_t)b.mGpu;
}

int DataX::UpdateGPUAccess() {
  // datax.cpp:81-94: publish the 16 device pointers.
  if (cuData == nullptr) return fx::fail(FX_ESTATE, "DataX::UpdateGPUAccess: not assigned to a symbol");
  cuDataX t;
  FillTable(&t);
  FX_HIP(hipMemcpyHtoD(cuData, &t, sizeof(cuDataX)), "DataX::UpdateGPUAccess");
  return FX_OK;
}

int DataX::MatchAllBuffers(DataX* src, uchar use_flags) {
  // datax.cpp:95-106
  DeleteAllBuffers();
  for (const auto& sb : src->mBuf) {
    FX_TRY(AddBuffer(sb.mRefID, "", (ushort)sb.mStride, sb.mMax,
                     use_flags != DT_MISC ? use_flags : sb.mUseFlags));
    SetBufferUsage(sb.mRefID, sb.mUseType);
  }
  return FX_OK;
}

int DataX::CopyAllBuffers(DataX* dest, uchar dest_flags, hipStream_t stream) {
  // datax.cpp:108-120; a layout mismatch is FX_EMISMATCH instead of exit(-9).
  if (dest->mBuf.size() != mBuf.size()) return fx::fail(FX_EMISMATCH, "DataX::CopyAllBuffers: buffer counts differ");
  for (size_t b = 0; b < mBuf.size(); b++) {
    if (mBuf[b].mRefID != dest->mBuf[b].mRefID)
      return fx::fail(FX_EMISMATCH, "DataX::CopyAllBuffers: RefIDs do not match");
    FX_TRY(mBuf[b].CopyTo(&dest->mBuf[b], dest_flags, stream));
  }
  return FX_OK;
}

int DataX::SwapGPU(DataX* other) {
  if (other->mBuf.size() != mBuf.size()) return fx::fail(FX_EMISMATCH, "DataX::SwapGPU: buffer counts differ");
  for (size_t b = 0; b < mBuf.size(); b++) {
    if (mBuf[b].mRefID != other->mBuf[b].mRefID || mBuf[b].mSize != other->mBuf[b].mSize)
      return fx::fail(FX_EMISMATCH, "DataX::SwapGPU: buffer %zu layouts differ", b);
  }
  for (size_t b = 0; b < mBuf.size(); b++) {
    void* t = mBuf[b].mGpu;
    mBuf[b].mGpu = other->mBuf[b].mGpu;
    other->mBuf[b].mGpu = t;
  }
  return FX_OK;
}

int DataX::Retrieve(int i, hipStream_t stream) {
  int b = mRef[i];
  if (b == BUNDEF) return FX_OK;
  return mBuf[b].Retrieve(stream);
}

int DataX::Commit(int i, hipStream_t stream) {
  int b = mRef[i];
  if (b == BUNDEF) return FX_OK;
  return mBuf[b].Commit(stream);
}

int DataX::CommitAll(hipStream_t stream) {
  for (auto& b : mBuf) FX_TRY(b.Commit(stream));
  return FX_OK;
}
