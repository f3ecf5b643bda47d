// cv_batch_legs.h -- the batched V-cycle below a chosen depth as persistent
// kernels: one launch per multigrid depth on the way down (pre-smoothing,
// residual and restriction of the cases that go deeper; coarse sweeps and
// the nullspace rule of the cases that bottom out there) and one per depth
// on the way up (prolongation and post-smoothing), for every case of the
// batch at once.  Included by cv_batch_solve.h.
//
// Work items.  A leg is a list of phases per case (a line sweep, a residual,
// a transfer, ...); a phase is split into items (a stencil tile, a block of
// y-lines, a chunk of points).  Work-groups of 64 x 16 threads take items in
// ticket order from one atomic counter.  The tickets run phase-major (every
// case's phase 0 items, then every case's phase 1 items, ...), and an item
// of phase p of case c waits until all items of phase p-1 of case c are
// done.  Every item a wait depends on holds a smaller ticket, and a
// work-group takes a ticket only while it runs and finishes that item before
// it takes another, so the earliest unfinished ticket can always proceed:
// the leg never deadlocks, whatever number of work-groups is resident.  A
// wait that runs past a spin limit (it cannot, short of a fault elsewhere)
// sets an error flag and every work-group leaves; the host reports it.
//
// Memory ordering across work-groups, which may sit on different XCDs with
// their own L2: each wave waits for its own stores, the work-group meets at
// a barrier, and one thread publishes the item with an agent-scope release
// add on the phase counter; the waiting thread polls with agent-scope loads
// and issues an agent-scope acquire fence before its work-group reads.
//
// Every item runs the single-case device body on the single-case operands:
// the stencil tile and line-solve results do not depend on the tile shape
// or the number of line chunks, elementwise passes do not depend on their
// partition, and the nullspace sum keeps the single-case partition of
// nred(n) blocks of BS threads (four such blocks per work-group).  A case
// therefore ends with the bits it gets alone.
#pragma once

#define LEG_TX 64
#define LEG_TY 16
#define LEG_NT (LEG_TX * LEG_TY)
#define LEG_CHUNK 4096                  // points per item of an elementwise phase
#define LEG_SPIN (1L << 22)             // polls before a wait gives up

enum { PK_LINE = 0, PK_STEN = 1, PK_RESTRICT = 2, PK_PROLONG = 3, PK_ZERO = 4, PK_FSUM = 5, PK_FFINAL = 6, PK_FSHIFT = 7 };

// one phase of one case in a leg
struct PPhase {
  int kind;      // PK_*
  int nitems;
  int src;       // PK_LINE: right-hand side slot (BB or BR)
  int zg;        // PK_LINE: zero initial guess
};

// a leg as its kernel sees it
struct PLeg {
  const uint32_t *ord;    // tickets: case << 26 | phase << 18 | item
  long ntk;
  const int *pbase;       // per case: first phase of the case in ph / done
  const PPhase *ph;
  unsigned *done;         // per phase: items finished
  unsigned *ticket;       // next ticket
};

#define LEG_CASE(e) ((int)((e) >> 26))
#define LEG_PHASE(e) ((int)(((e) >> 18) & 255u))
#define LEG_ITEM(e) ((int)((e) & 0x3ffffu))

template <class T, bool HG>
__global__ void __launch_bounds__(LEG_NT) kb_leg(const BCase *__restrict__ cs, unsigned long long pm,
                                                const BCtl *__restrict__ ctl, PLeg lg, int d, T omega, int kc,
                                                long spin, int *err) {
  __shared__ uint32_t s_tk;
  __shared__ int s_ok;
  __shared__ double shv[LEG_NT / 64];
  const int tid = threadIdx.y * LEG_TX + threadIdx.x;
  const unsigned long long on = pm & ctl->act;
  for (;;) {
    __syncthreads();   // every thread has read the previous ticket
    if (tid == 0) s_tk = __hip_atomic_fetch_add(lg.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tk = s_tk;
    if ((long)tk >= lg.ntk) break;
    const uint32_t e = lg.ord[tk];
    const int c = LEG_CASE(e), p = LEG_PHASE(e), it = LEG_ITEM(e);
    // a case outside the mask neither waits nor signals: all its items skip
    if (!((on >> c) & 1ull)) continue;
    const int q = lg.pbase[c] + p;
    if (p > 0) {
      if (tid == 0) {
        const unsigned need = (unsigned)lg.ph[q - 1].nitems;
        // spin < 0: every wait fails at once (the error path, forced from the host)
        int ok = spin >= 0;
        long spins = 0;
        while (ok && __hip_atomic_load(lg.done + q - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++spins > spin || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        if (!ok) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
      }
      __syncthreads();
      if (!s_ok) break;
    }
    const BCase &C = cs[c];
    const PPhase P = lg.ph[q];
    const BLev &Lv = C.lev[d];
    const OpV<T> A = bop<T>(C, d);
    T *x = bvec<T>(C, d, BX, ctl->jj);
    const long c0 = (long)it * LEG_CHUNK, c1 = c0 + LEG_CHUNK;
    switch (P.kind) {
      case PK_STEN: {
        const int gx = (A.mx + LEG_TX - 2) / (LEG_TX - 1), gy = (A.my + LEG_TY - 1) / LEG_TY;
        stencil_tile<T, 1, LEG_TY, HG, HG, LEG_TX>(A, x, bvec<T>(C, d, BB, 0), bvec<T>(C, d, BR, 0), A.dinv, (T)0,
                                                     kc, it % gx, (it / gx) % gy, it / (gx * gy));
        break;
      }
      case PK_LINE: {
        const T om = HG ? (T)(omega * A.ysc) : omega;
        const T *r = bvec<T>(C, d, P.src, 0);
        if (Lv.cs == 8)
          yline_reg_tile<T, 8, HG>(A.mx, A.my, A.mz, r, A.ylo, A.yip, A.ycp, x, x, om, P.zg, it, A.yh, A.yc);
        else
          yline_reg_tile<T, 16, HG>(A.mx, A.my, A.mz, r, A.ylo, A.yip, A.ycp, x, x, om, P.zg, it, A.yh, A.yc);
        break;
      }
      case PK_RESTRICT: {
        const BLev &K = C.lev[d + 1];
        restrict_body<T>(K.I, Lv.I, K.isc, K.jsc, K.ksc, Lv.nv, Lv.nwf, K.fl, (const T *)bvec<T>(C, d, BR, 0),
                         bvec<T>(C, d + 1, BB, 0), c1 < K.n ? c1 : K.n, c0 + tid, LEG_NT);
        break;
      }
      case PK_PROLONG: {
        const BLev &K = C.lev[d + 1];
        prolong_body<T>(Lv.I, K.I, K.isc, K.jsc, K.ksc, K.nv, Lv.fl, (const T *)bvec<T>(C, d + 1, BX, ctl->jj), x, 1,
                        c1 < Lv.n ? c1 : Lv.n, c0 + tid, LEG_NT);
        break;
      }
      case PK_ZERO: {
        const long n1 = c1 < Lv.n ? c1 : Lv.n;
        for (long t = c0 + tid; t < n1; t += LEG_NT) x[t] = (T)0;
        break;
      }
      case PK_FSUM: {
        // four reduction blocks of BS threads, each as the single-case launch
        const int g = tid / BS, lt = tid % BS, vb = it * (LEG_NT / BS) + g;
        double s = vb < Lv.nrd ? fluid_sum_thread<T>(x, Lv.fl, Lv.n, vb, Lv.nrd, lt) : 0.0;
        s = cv_wave_sum(s);
        if ((tid & 63) == 0) shv[tid >> 6] = s;
        __syncthreads();
        if (lt == 0 && vb < Lv.nrd) {
          double a = 0.0;
          for (int w = 0; w < BS / 64; w++) a += shv[g * (BS / 64) + w];
          C.part[vb] = a;
        }
        break;
      }
      case PK_FFINAL: {
        // the single-case final pass: one block of BS threads over the partials
        double s = 0.0;
        if (tid < BS)
          for (int b = tid; b < Lv.nrd; b += BS) s += C.part[b];
        s = cv_wave_sum(s);
        if ((tid & 63) == 0) shv[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
          double a = 0.0;
          for (int w = 0; w < BS / 64; w++) a += shv[w];
          C.dsc[100] = a;
        }
        break;
      }
      default: {   // PK_FSHIFT
        fluid_shift_body<T>(x, Lv.fl, C.dsc + 100, Lv.inv_nf, c1 < Lv.n ? c1 : Lv.n, c0 + tid, LEG_NT);
        break;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have completed
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(lg.done + q, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------- host side -------------------------------------------
struct BLegHost {
  int depth = 0, up = 0;
  long off_ord = 0, ntk = 0;      // into ord_all
  long off_ph = 0, nph = 0;       // into ph_all / done_all; phases of the leg
  long off_pb = 0;                // into pbase_all (n entries)
  int ticket = 0;                 // index of its ticket counter in done_all (after the phases)
  unsigned long long mask = 0;    // cases with work in the leg
};

struct BLegs {
  int d0 = 0;                     // first depth run by legs
  int kc = 4;                     // k planes per stencil item
  long spin = LEG_SPIN;           // polls before a wait gives up (negative: fail every wait)
  int grid = 0;                   // work-groups per leg launch
  std::vector<BLegHost> down, up; // by depth (down[d - d0], up[d - d0])
  uint32_t *ord = nullptr;
  PPhase *ph = nullptr;
  int *pbase = nullptr;
  unsigned *done = nullptr;       // phase counters of every leg, then one ticket per leg
  long ndone = 0;
  int *err = nullptr;
};

static void legs_free(BLegs *L) {
  if (!L) return;
  if (L->ord) (void)hipFree(L->ord);
  if (L->ph) (void)hipFree(L->ph);
  if (L->pbase) (void)hipFree(L->pbase);
  if (L->done) (void)hipFree(L->done);
  if (L->err) (void)hipFree(L->err);
  delete L;
}

// items of a phase of case c at depth d
static int leg_items(const BCase &C, int d, int kind, int kc) {
  const BLev &b = C.lev[d];
  switch (kind) {
    case PK_STEN: {
      const int gx = (b.I.mx + LEG_TX - 2) / (LEG_TX - 1), gy = (b.I.my + LEG_TY - 1) / LEG_TY, gz = (b.I.mz + kc - 1) / kc;
      return gx * gy * gz;
    }
    case PK_LINE: return b.lgx;
    case PK_RESTRICT: return (int)((C.lev[d + 1].n + LEG_CHUNK - 1) / LEG_CHUNK);
    case PK_FSUM: return (b.nrd + LEG_NT / BS - 1) / (LEG_NT / BS);
    case PK_FFINAL: return 1;
    default: return (int)((b.n + LEG_CHUNK - 1) / LEG_CHUNK);
  }
}

// the phases of case c in the down (up = 0) or up leg of depth d; the same
// operation sequence as b_vcycle for that case
static void leg_phases(const BCase &C, int d, int up, const cv_solver_opts &o, int kc, std::vector<PPhase> &ph) {
  auto add = [&](int kind, int src = BB, int zg = 0) { ph.push_back(PPhase{kind, leg_items(C, d, kind, kc), src, zg}); };
  const bool coarse = d == C.nlev - 1;
  if (up) {
    add(PK_PROLONG);
    for (int s = 0; s < o.post_sweeps; s++) { add(PK_STEN); add(PK_LINE, BR, 0); }
    return;
  }
  const int nsw = coarse ? o.coarse_sweeps : o.pre_sweeps;
  if (nsw == 0) add(PK_ZERO);
  for (int s = 0; s < nsw; s++) {
    if (s > 0) add(PK_STEN);
    add(PK_LINE, s == 0 ? BB : BR, s == 0);
  }
  if (coarse) {
    add(PK_FSUM);
    add(PK_FFINAL);
    add(PK_FSHIFT);
  } else {
    add(PK_STEN);
    add(PK_RESTRICT);
  }
}

// the legs from depth d0 down; returns CV_OK, or an error with nothing kept
static int legs_build(cv_batch *B, const std::vector<BCase> &hc, int d0, int kc, BLegs **out) {
  *out = nullptr;
  const int n = B->n;
  // one leg launch fills every CU with as many work-groups as fit at once
  int cus = 0, per = 0;
  CV_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->device));
  if (B->opt.precision == CV_MIXED16)
    CV_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kb_leg<float, true>, LEG_NT, 0));
  else if (B->opt.precision == CV_MIXED)
    CV_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kb_leg<float, false>, LEG_NT, 0));
  else
    CV_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kb_leg<double, false>, LEG_NT, 0));
  const int grid = std::max(cus, 1) * std::max(per, 1);
  BLegs *L = new BLegs();
  L->d0 = d0;
  L->kc = kc;
  const char *es = getenv("CV_BATCH_LEG_SPIN");
  if (es) L->spin = atol(es);
  std::vector<uint32_t> ord;
  std::vector<PPhase> ph;
  std::vector<int> pbase;
  long nph = 0;
  std::vector<BLegHost *> all;
  for (int d = d0; d < B->maxlev; d++) {
    L->down.emplace_back();
    L->up.emplace_back();
  }
  for (int d = d0; d < B->maxlev; d++)
    for (int up = 0; up < 2; up++) {
      BLegHost &h = up ? L->up[d - d0] : L->down[d - d0];
      h.depth = d;
      h.up = up;
      h.off_ord = (long)ord.size();
      h.off_pb = (long)pbase.size();
      h.off_ph = nph;
      std::vector<std::vector<PPhase>> per(n);
      int pmax = 0;
      for (int c = 0; c < n; c++) {
        pbase.push_back(0);
        const bool has = hc[c].nlev > d && (!up || d < hc[c].nlev - 1);
        if (!has) continue;
        leg_phases(hc[c], d, up, B->opt, kc, per[c]);
        pmax = std::max(pmax, (int)per[c].size());
        h.mask |= 1ull << c;
      }
      for (int c = 0; c < n; c++) {
        pbase[h.off_pb + c] = (int)(nph - h.off_ph);
        for (const PPhase &p : per[c]) {
          if (p.nitems <= 0 || p.nitems >= (1 << 18) || per[c].size() > 255) {
            legs_free(L);
            cv_set_error("cv_batch_create: a persistent V-cycle phase is out of range");
            return CV_ERR_ARG;
          }
          ph.push_back(p);
          nph++;
        }
      }
      // phase-major ticket order
      for (int p = 0; p < pmax; p++)
        for (int c = 0; c < n; c++)
          if (p < (int)per[c].size())
            for (int i = 0; i < per[c][p].nitems; i++)
              ord.push_back(((uint32_t)c << 26) | ((uint32_t)p << 18) | (uint32_t)i);
      h.ntk = (long)ord.size() - h.off_ord;
      h.nph = nph - h.off_ph;
    }
  // host check of the tables the kernel trusts: every ticket names a phase of
  // its case, and an item below that phase's count; the phases of a case
  // appear in order, so each wait names a phase with smaller tickets only
  for (int q = 0; q < 2 * (B->maxlev - d0); q++) {
    const BLegHost &h = (q & 1) ? L->up[q / 2] : L->down[q / 2];
    std::vector<int> last(n, -1);
    for (long t = 0; t < h.ntk; t++) {
      const uint32_t e = ord[h.off_ord + t];
      const int c = LEG_CASE(e), p = LEG_PHASE(e), it = LEG_ITEM(e);
      const long pq = h.off_ph + pbase[h.off_pb + c] + p;
      const bool okc = c < n && ((h.mask >> c) & 1ull);
      const long pend = !okc ? 0 : c + 1 < n ? h.off_ph + pbase[h.off_pb + c + 1] : h.off_ph + h.nph;
      const bool ok = okc && pq < pend && it < ph[pq].nitems && p >= last[c];
      if (!ok) {
        legs_free(L);
        cv_set_error("cv_batch_create: inconsistent persistent V-cycle tables");
        return CV_ERR_ARG;
      }
      last[c] = p;
    }
  }
  const int nlegs = 2 * (B->maxlev - d0);
  L->ndone = nph + nlegs;
  for (int q = 0; q < nlegs; q++) {
    BLegHost &h = (q & 1) ? L->up[q / 2] : L->down[q / 2];
    h.ticket = (int)(nph + q);
  }
  L->grid = grid;
  bool okm = hipMalloc((void **)&L->ord, sizeof(uint32_t) * std::max<size_t>(ord.size(), 1)) == hipSuccess &&
             hipMalloc((void **)&L->ph, sizeof(PPhase) * std::max<size_t>(ph.size(), 1)) == hipSuccess &&
             hipMalloc((void **)&L->pbase, sizeof(int) * std::max<size_t>(pbase.size(), 1)) == hipSuccess &&
             hipMalloc((void **)&L->done, sizeof(unsigned) * L->ndone) == hipSuccess &&
             hipMalloc((void **)&L->err, sizeof(int)) == hipSuccess;
  if (!okm) {
    legs_free(L);
    cv_set_error("cv_batch_create: out of device memory");
    return CV_ERR_OOM;
  }
  hipError_t he = hipSuccess;
  if (he == hipSuccess && !ord.empty()) he = hipMemcpy(L->ord, ord.data(), sizeof(uint32_t) * ord.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess && !ph.empty()) he = hipMemcpy(L->ph, ph.data(), sizeof(PPhase) * ph.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess && !pbase.empty())
    he = hipMemcpy(L->pbase, pbase.data(), sizeof(int) * pbase.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemset(L->done, 0, sizeof(unsigned) * L->ndone);
  if (he == hipSuccess) he = hipMemset(L->err, 0, sizeof(int));
  if (he != hipSuccess) {
    legs_free(L);
    cv_set_error("cv_batch_create: %s", hipGetErrorString(he));
    return CV_ERR_HIP;
  }
  *out = L;
  return CV_OK;
}

template <class T>
static void leg_launch(cv_batch *B, const BLegHost &h, unsigned long long pm) {
  const BLegs *L = B->legs;
  if (h.ntk == 0 || !(pm & h.mask)) return;
  PLeg lg;
  lg.ord = L->ord + h.off_ord;
  lg.ntk = h.ntk;
  lg.pbase = L->pbase + h.off_pb;
  lg.ph = L->ph + h.off_ph;
  lg.done = L->done + h.off_ph;
  lg.ticket = L->done + h.ticket;
  const unsigned grid = (unsigned)std::min<long>(h.ntk, L->grid);
  const T om = (T)B->opt.omega;
  if (sizeof(T) == 4 && B->opt.precision == CV_MIXED16)
    hipLaunchKernelGGL((kb_leg<T, sizeof(T) == 4>), dim3(grid), dim3(LEG_TX, LEG_TY), 0, B->stream,
                       (const BCase *)B->dcases, pm, (const BCtl *)B->dctl, lg, h.depth, om, L->kc, L->spin, L->err);
  else
    hipLaunchKernelGGL((kb_leg<T, false>), dim3(grid), dim3(LEG_TX, LEG_TY), 0, B->stream, (const BCase *)B->dcases, pm,
                       (const BCtl *)B->dctl, lg, h.depth, om, L->kc, L->spin, L->err);
}

// the V-cycle from depth d0 down and back, for every case of the batch: the
// counters start at zero (a kernel, not a memset: this runs inside a graph)
template <class T>
static void b_vcycle_legs(cv_batch *B) {
  BLegs *L = B->legs;
  hipLaunchKernelGGL(k_zero<unsigned>, dim3((unsigned)std::min<long>((L->ndone + BS - 1) / BS, 64)), dim3(BS), 0,
                     B->stream, L->done, L->ndone);
  const int nd = B->maxlev - L->d0;
  for (int q = 0; q < nd; q++) leg_launch<T>(B, L->down[q], B->all);
  for (int q = nd - 2; q >= 0; q--) leg_launch<T>(B, L->up[q], B->all);
}

// queue the read-back of the legs' error flag (b_gather checks it after its sync)
static int b_legs_error(cv_batch *B) {
  if (!B->legs) return CV_OK;
  CV_CHECK(hipMemcpyAsync(B->herr, B->legs->err, sizeof(int), hipMemcpyDeviceToHost, B->stream));
  return CV_OK;
}
