// Kernels as argument structs, launched one problem at a time or many
// problems in one launch.
//
// A kernel is a struct K holding its arguments, with
//   static constexpr int kThreads;                 // threads per workgroup
//   __device__ void run(int block, int nblocks) const;
// run() is the body of one workgroup; block / nblocks replace blockIdx.x /
// gridDim.x.  launch(k, nblocks, s) runs it as an ordinary launch
// (k_single).  The same struct, for n problems, runs as ONE launch (k_multi):
// workgroup b belongs to problem p = pmap[b] and runs as block b - off[p] of
// off[p + 1] - off[p], so every problem's workgroups do exactly the work of
// its own launch, in the same order, and its results are bitwise those of
// its own launch.
//
// A Recorder, while installed on the calling thread, keeps a copy of every
// launch made through launch() (the launches still happen); the batch solve
// uses this to collect the launch sequence of each problem's PCG iteration
// and to merge equal sequences into one launch per step.
#pragma once
#include <memory>
#include <typeinfo>
#include <vector>

#include "common.h"

namespace gmgs {

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_single(const K k) {
  k.run((int)blockIdx.x, (int)gridDim.x);
}

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_multi(const K* __restrict__ ks, const int* __restrict__ pmap,
                                                       const int* __restrict__ off) {
  const int b = (int)blockIdx.x;
  // (uniform: keeps the argument loads scalar)
  const int p = __builtin_amdgcn_readfirstlane(pmap[b]);
  const int o0 = __builtin_amdgcn_readfirstlane(off[p]), o1 = __builtin_amdgcn_readfirstlane(off[p + 1]);
  ks[p].run(b - o0, o1 - o0);
}

// one recorded launch: the kernel's type, its workgroups and a copy of its
// argument struct
struct LaunchRec {
  const std::type_info* type = nullptr;
  long long nblocks = 0;
  std::shared_ptr<const void> args;
  // uploads the argument structs of recs (all of this record's type) and
  // the block maps into 'keep', and returns a closure that launches them as
  // one k_multi on a stream
  struct Multi {
    long long total = 0;
    const void* ks = nullptr;
    const int* pmap = nullptr;
    const int* off = nullptr;
    void (*go)(const Multi&, hipStream_t) = nullptr;
  };
  Multi (*prepare)(const std::vector<const LaunchRec*>& recs, std::vector<DevBuf>& keep,
                   hipStream_t s) = nullptr;
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

// the recorder installed on this thread, or null
Recorder*& current_recorder();

template <class K>
LaunchRec::Multi prepare_multi(const std::vector<const LaunchRec*>& recs, std::vector<DevBuf>& keep,
                               hipStream_t s) {
  std::vector<K> ks;
  std::vector<int> off(1, 0), pmap;
  ks.reserve(recs.size());
  for (size_t p = 0; p < recs.size(); ++p) {
    const LaunchRec* r = recs[p];
    if (*r->type != typeid(K)) throw std::logic_error("batched launch: kernels differ between problems");
    ks.push_back(*static_cast<const K*>(r->args.get()));
    if (off.back() + r->nblocks >= (1LL << 31)) throw std::invalid_argument("batched launch too large");
    off.push_back(off.back() + (int)r->nblocks);
    pmap.insert(pmap.end(), (size_t)r->nblocks, (int)p);
  }
  LaunchRec::Multi m;
  m.total = off.back();
  if (m.total == 0) return m;
  DevBuf kb, pb, ob;
  kb.alloc(sizeof(K) * ks.size());
  pb.alloc(sizeof(int) * pmap.size());
  ob.alloc(sizeof(int) * off.size());
  GMGS_HIP(hipMemcpyAsync(kb.as<void>(), ks.data(), sizeof(K) * ks.size(), hipMemcpyHostToDevice, s));
  GMGS_HIP(hipMemcpyAsync(pb.as<void>(), pmap.data(), sizeof(int) * pmap.size(), hipMemcpyHostToDevice, s));
  GMGS_HIP(hipMemcpyAsync(ob.as<void>(), off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, s));
  GMGS_HIP(hipStreamSynchronize(s));   // the host vectors go out of scope
  m.ks = kb.as<void>();
  m.pmap = pb.as<int>();
  m.off = ob.as<int>();
  m.go = [](const LaunchRec::Multi& mm, hipStream_t st) {
    k_multi<K><<<(unsigned)mm.total, K::kThreads, 0, st>>>(static_cast<const K*>(mm.ks), mm.pmap, mm.off);
    GMGS_LAUNCH_CHECK();
  };
  keep.push_back(std::move(kb));
  keep.push_back(std::move(pb));
  keep.push_back(std::move(ob));
  return m;
}

// run k on nblocks workgroups of K::kThreads threads (no launch for none)
template <class K>
void launch(const K& k, long long nblocks, hipStream_t s) {
  if (nblocks <= 0) return;
  if (nblocks >= (1LL << 31)) throw std::invalid_argument("grid too large");
  if (Recorder* r = current_recorder()) {
    LaunchRec rec;
    rec.type = &typeid(K);
    rec.nblocks = nblocks;
    rec.args = std::make_shared<const K>(k);
    rec.prepare = &prepare_multi<K>;
    r->recs.push_back(std::move(rec));
  }
  k_single<K><<<(unsigned)nblocks, K::kThreads, 0, s>>>(k);
  GMGS_LAUNCH_CHECK();
}

}  // namespace gmgs
