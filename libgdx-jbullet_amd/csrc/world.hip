// world.hip — libjbgpu host side: the C ABI of include/jbgpu.h, the
// device-resident step (DiscreteDynamicsWorld.stepSimulation with one substep,
// dynamics/DiscreteDynamicsWorld.java:290-400) and the per-body kernels:
// integrateVelocities/applyDamping (dynamics/RigidBody.java:246-342),
// integrateTransform (linearmath/TransformUtil.java:64-103) and the
// activation update (DiscreteDynamicsWorld.java:475-512, RigidBody.java:543-576).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "../../include/jbgpu.h"
#include "internal.h"
#include "mlp.h"
#include "shapes.h"

using namespace jbg;

struct jbgpu_ctx {
  Ctx c;
};

namespace {

std::string g_createError;

constexpr float MAX_ANGVEL = kHalfPi;

__global__ void kPredict(int n, float dt, const float* __restrict__ tr, const float* __restrict__ props,
                         const float* __restrict__ grav, const int* __restrict__ act, float* vel,
                         const int* __restrict__ abortFlag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (abortFlag && *abortFlag)) return;  // an aborted graph-captured step changes nothing
  const float* pr = props + 8 * i;
  if ((ldi(pr + 7) & JBGPU_BODY_STATIC) != 0) return;
  int a = act[i];
  if (a == JBGPU_ISLAND_SLEEPING || a == JBGPU_DISABLE_SIMULATION) return;
  float* v = vel + 8 * i;
  V3 lv = ld3(v), av = ld3(v + 4);
  float im = pr[0];
  // applyGravity: totalForce = 0 + gravity * linearFactor
  V3 g = ld3(grav + 4 * i);
  V3 tf = v3(0.0f + g.x * 1.0f, 0.0f + g.y * 1.0f, 0.0f + g.z * 1.0f);
  lv.x += im * dt * tf.x;
  lv.y += im * dt * tf.y;
  lv.z += im * dt * tf.z;
  // angular: torque 0 through the world inverse inertia (RigidBody.java:331-334)
  M3 b;
  for (int k = 0; k < 9; k++) b.m[k] = tr[12 * i + k];
  M3 m1;
  for (int r = 0; r < 3; r++) {
    m1.m[r * 3 + 0] = b.m[r * 3 + 0] * pr[1];
    m1.m[r * 3 + 1] = b.m[r * 3 + 1] * pr[2];
    m1.m[r * 3 + 2] = b.m[r * 3 + 2] * pr[3];
  }
  M3 iw = mulMM(m1, transpose(b));
  V3 t = scl(mulMV(iw, v3(0, 0, 0)), dt);
  av = add(av, t);
  float angvel = len(av);
  if (angvel * dt > MAX_ANGVEL) av = scl(av, (MAX_ANGVEL / dt) / angvel);
  float ld = (float)pow(1.0 - 0.0, (double)dt);  // applyDamping with zero damping
  lv = scl(lv, ld);
  av = scl(av, ld);
  st3(v, lv);
  st3(v + 4, av);
}

__global__ void kIntegrate(int n, float dt, float* tr, const float* __restrict__ props, const int* __restrict__ act,
                           const float* __restrict__ vel, const int* __restrict__ abortFlag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (abortFlag && *abortFlag)) return;
  const float* pr = props + 8 * i;
  if ((ldi(pr + 7) & JBGPU_BODY_STATIC) != 0) return;
  int a = act[i];
  if (a == JBGPU_ISLAND_SLEEPING || a == JBGPU_DISABLE_SIMULATION) return;
  float* T = tr + 12 * i;
  Tr t;
  for (int k = 0; k < 9; k++) t.b.m[k] = T[k];
  t.o = v3(T[9], T[10], T[11]);
  const Tr r = integrateTr(t, ld3(vel + 8 * i), ld3(vel + 8 * i + 4), dt);
  for (int k = 0; k < 9; k++) T[k] = r.b.m[k];
  T[9] = r.o.x;
  T[10] = r.o.y;
  T[11] = r.o.z;
}

// plain device copy of a known size (bandwidth ceiling and counter scale reference)
__global__ __launch_bounds__(256) void kCopyProbe(int64_t n4, const float4* __restrict__ src, float4* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void kActivation(int n, float dt, int deactDisabled, const float* __restrict__ props, int* act,
                            float* deact, float* vel, const int* __restrict__ abortFlag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (abortFlag && *abortFlag)) return;
  int a = act[i];
  float t = deact[i];
  float* v = vel + 8 * i;
  auto setAct = [&](int s) {
    if (a != JBGPU_DISABLE_DEACTIVATION && a != JBGPU_DISABLE_SIMULATION) a = s;
  };
  if (!(a == JBGPU_ISLAND_SLEEPING || a == JBGPU_DISABLE_DEACTIVATION)) {
    if ((len2(ld3(v)) < 0.8f * 0.8f) && (len2(ld3(v + 4)) < 1.0f * 1.0f)) {
      t += dt;
    } else {
      t = 0.0f;
      setAct(0);
    }
  }
  bool wants;
  if (a == JBGPU_DISABLE_DEACTIVATION) wants = false;
  else if (deactDisabled) wants = false;
  else if (a == JBGPU_ISLAND_SLEEPING || a == JBGPU_WANTS_DEACTIVATION) wants = true;
  else wants = t > 2.0f;
  bool isStatic = (ldi(props + 8 * i + 7) & JBGPU_BODY_STATIC) != 0;
  if (wants) {
    if (isStatic) {
      setAct(JBGPU_ISLAND_SLEEPING);
    } else {
      if (a == JBGPU_ACTIVE_TAG) setAct(JBGPU_WANTS_DEACTIVATION);
      if (a == JBGPU_ISLAND_SLEEPING) {
        st3(v, v3(0, 0, 0));
        st3(v + 4, v3(0, 0, 0));
      }
    }
  } else {
    if (a != JBGPU_DISABLE_DEACTIVATION) setAct(JBGPU_ACTIVE_TAG);
  }
  act[i] = a;
  deact[i] = t;
}

template <class F>
int guard(jbgpu_ctx* ctx, F f) {
  if (!ctx) return JBGPU_ERR_ARG;
  try {
    f();
    return JBGPU_OK;
  } catch (const std::invalid_argument& e) {
    ctx->c.err = e.what();
    return JBGPU_ERR_ARG;
  } catch (const std::length_error& e) {
    ctx->c.err = e.what();
    return JBGPU_ERR_CAPACITY;
  } catch (const std::logic_error& e) {
    ctx->c.err = e.what();
    return JBGPU_ERR_STATE;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return JBGPU_ERR_HIP;
  }
}

void requireWorld(const Ctx& c) {
  if (!c.haveWorld) throw std::logic_error("no world: call jbgpu_world_create first");
}

template <class T>
void upload(DBuf<T>& b, const T* src, size_t n, hipStream_t s) {
  b.ensure(std::max<size_t>(n, 1));
  if (n) JB_CHECK(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

// proxy state for n proxies, keeping the first c.np (appended proxies follow the existing ones)
void allocProxyState(Ctx& c, int n) {
  c.stage.ensureKeep(n);
  c.leaf.ensureKeep(6 * (size_t)n);
  c.aabb.ensureKeep(6 * (size_t)n);
  c.q.ensureKeep(6 * (size_t)n);
  c.filt.ensureKeep(n);
  c.world.ensureKeep(n);
  c.proxyDead.resize(n, 0);
  c.np = n;
}

// createProxy for n new proxies after the existing ones (uids np + 1 ..); `replace` starts from none
void uploadProxies(Ctx& c, int n, const float* aabb6, const int16_t* group, const int16_t* mask, const int32_t* world,
                   bool replace) {
  if (replace) {
    c.np = 0;
    c.maxWorld = 0;
    c.proxyDead.clear();
    c.npairs = 0;
    c.npairsOld = 0;
  }
  const int n0 = c.np;
  if ((int64_t)n0 + n >= (1ll << 30)) throw std::length_error("too many proxies");
  std::vector<int> stage(n, c.stageCurrent), filt(n), wid(n);
  std::vector<int> q(6 * (size_t)n, 0);
  std::vector<float> box(aabb6, aabb6 + 6 * (size_t)n);
  for (int i = 0; i < n; i++) {
    int16_t g = group ? group[i] : (int16_t)1, m = mask ? mask[i] : (int16_t)-1;
    filt[i] = (int)(uint16_t)g | ((int)(uint16_t)m << 16);
    wid[i] = world ? world[i] : 0;
    if (wid[i] < 0) throw std::invalid_argument("negative world id");
  }
  allocProxyState(c, n0 + n);
  for (int i = 0; i < n; i++) c.maxWorld = std::max(c.maxWorld, wid[i]);
  c.worldBits = 0;
  while (c.worldBits < 31 && (1ll << c.worldBits) <= c.maxWorld) c.worldBits++;
  if (c.bpMode == JBGPU_BP_SAP16 || c.bpMode == JBGPU_BP_SAP32) {
    // AxisSweep3Internal.quantize (host copy of the device routine)
    for (int i = 0; i < n; i++)
      for (int mx = 0; mx < 2; mx++)
        for (int k = 0; k < 3; k++) {
          float p = box[6 * i + 3 * mx + k];
          p = jmax(p, c.wmin[k]);
          p = jmin(p, c.wmax[k]);
          float v = (p - c.wmin[k]) * c.quant[k];
          q[6 * i + 3 * mx + k] = (int)((((uint32_t)jint(v) & c.handleMask) | (uint32_t)mx) & c.qmask);
        }
  }
  hipStream_t s = c.stream;
  JB_CHECK(hipMemcpyAsync(c.stage.p + n0, stage.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.leaf.p + 6 * (size_t)n0, box.data(), 6 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.aabb.p + 6 * (size_t)n0, box.data(), 6 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.q.p + 6 * (size_t)n0, q.data(), 6 * (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.filt.p + n0, filt.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.world.p + n0, wid.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipStreamSynchronize(s));
}

void bodiesToDevice(Ctx& c, const float* bodies) {
  int n = c.nb;
  std::vector<float> tr(12 * (size_t)n), vel(8 * (size_t)n, 0.0f), props(8 * (size_t)n), grav(4 * (size_t)n, 0.0f),
      deact(n);
  std::vector<int> act(n);
  for (int i = 0; i < n; i++) {
    const float* r = bodies + (size_t)JBGPU_BODY_WORDS * i;
    for (int k = 0; k < 12; k++) tr[12 * i + k] = r[k];
    for (int k = 0; k < 3; k++) {
      vel[8 * i + k] = r[12 + k];
      vel[8 * i + 4 + k] = r[15 + k];
      grav[4 * i + k] = r[22 + k];
    }
    props[8 * i + 0] = r[18];
    props[8 * i + 1] = r[19];
    props[8 * i + 2] = r[20];
    props[8 * i + 3] = r[21];
    props[8 * i + 4] = r[25];
    props[8 * i + 5] = r[26];
    props[8 * i + 6] = r[27];
    props[8 * i + 7] = r[28];
    act[i] = ldi(r + 29);
    deact[i] = r[30];
    int si = ldi(r + 27);
    if (si < 0 || si >= c.ns) throw std::invalid_argument("body shape index out of range");
  }
  c.bodyShapeType.resize(n);
  c.bodyStaticHost.resize(n);
  c.bodyInvMassHost.resize(n);
  for (int i = 0; i < n; i++) {
    const float* r = bodies + (size_t)JBGPU_BODY_WORDS * i;
    c.bodyShapeType[i] = c.shapeTypeHost[ldi(r + 27)];
    c.bodyStaticHost[i] = (ldi(r + 28) & JBGPU_BODY_STATIC) != 0;
    c.bodyInvMassHost[i] = r[18];
  }
  c.wbReady = false;
  jointsChanged(c);
  hipStream_t s = c.stream;
  upload(c.tr, tr.data(), tr.size(), s);
  upload(c.vel, vel.data(), vel.size(), s);
  upload(c.props, props.data(), props.size(), s);
  upload(c.grav, grav.data(), grav.size(), s);
  upload(c.act, act.data(), act.size(), s);
  upload(c.deact, deact.data(), deact.size(), s);
  JB_CHECK(hipStreamSynchronize(s));
}

void bodiesFromDevice(Ctx& c, float* bodies) {
  int n = c.nb;
  std::vector<float> tr(12 * (size_t)n), vel(8 * (size_t)n), props(8 * (size_t)n), grav(4 * (size_t)n), deact(n);
  std::vector<int> act(n);
  hipStream_t s = c.stream;
  JB_CHECK(hipMemcpyAsync(tr.data(), c.tr.p, tr.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(vel.data(), c.vel.p, vel.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(props.data(), c.props.p, props.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(grav.data(), c.grav.p, grav.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(act.data(), c.act.p, act.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(deact.data(), c.deact.p, deact.size() * 4, hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < n; i++) {
    float* r = bodies + (size_t)JBGPU_BODY_WORDS * i;
    std::memset(r, 0, JBGPU_BODY_WORDS * sizeof(float));
    for (int k = 0; k < 12; k++) r[k] = tr[12 * i + k];
    for (int k = 0; k < 3; k++) {
      r[12 + k] = vel[8 * i + k];
      r[15 + k] = vel[8 * i + 4 + k];
      r[22 + k] = grav[4 * i + k];
    }
    r[18] = props[8 * i + 0];
    r[19] = props[8 * i + 1];
    r[20] = props[8 * i + 2];
    r[21] = props[8 * i + 3];
    r[25] = props[8 * i + 4];
    r[26] = props[8 * i + 5];
    r[27] = props[8 * i + 6];
    r[28] = props[8 * i + 7];
    sti(r + 29, act[i]);
    r[30] = deact[i];
  }
}

// Host check of caller-supplied pair keys and manifold records before any of them reaches a kernel: keys
// strictly increasing with 1 <= uid_lo < uid_hi <= nbodies; per record a point count in [0, 4], a pair type in
// [0, 3] and the key's two bodies (in either order: a convex-plane record lists the convex body first).
void checkPairRecords(int nbodies, int64_t n, const uint64_t* keys, const float* man) {
  if (n < 0) throw std::invalid_argument("negative pair count");
  if (n == 0) return;
  if (!keys || !man) throw std::invalid_argument("null pair arrays");
  for (int64_t i = 0; i < n; i++) {
    const uint64_t k = keys[i];
    const int64_t lo = (int64_t)(k >> 32), hi = (int64_t)(k & 0xffffffffu);
    if (lo < 1 || hi <= lo || hi > nbodies)
      throw std::invalid_argument("pair " + std::to_string(i) + ": key uids out of range or not uid_lo < uid_hi");
    if (i > 0 && keys[i - 1] >= k) throw std::invalid_argument("pair keys not strictly increasing at " + std::to_string(i));
    const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * i;
    const int cnt = ldi(m), b0 = ldi(m + 1), b1 = ldi(m + 2), type = ldi(m + 3);
    if (cnt < 0 || cnt > 4) throw std::invalid_argument("pair " + std::to_string(i) + ": point count not in [0, 4]");
    if (type < 0 || type > JBGPU_PAIR_NONE)
      throw std::invalid_argument("pair " + std::to_string(i) + ": pair type not in [0, 3]");
    const int a = (int)lo - 1, b = (int)hi - 1;
    if (!((b0 == a && b1 == b) || (b0 == b && b1 == a)))
      throw std::invalid_argument("pair " + std::to_string(i) + ": manifold bodies do not match the key");
  }
}

// The manifold header the dispatcher keeps for a pair (count, body0, body1, type): bodies in uid order, except a
// convex-plane pair whose lower uid is the plane, which lists the convex body first (the device's remap and the
// oracle build the same header).  Records of a caller's pair set for the narrowphase seam must carry it: a
// record with points gets ERR_ARG otherwise (its points' A/B sides would be swapped); a record without points
// gets it written here.
void canonicalHeaders(const Ctx& c, int64_t n, const uint64_t* keys, float* man) {
  for (int64_t i = 0; i < n; i++) {
    const int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
    const int ta = c.bodyShapeType[a], tb = c.bodyShapeType[b];
    int type;
    if (ta == JBGPU_SHAPE_SPHERE && tb == JBGPU_SHAPE_SPHERE) type = JBGPU_PAIR_SPHERE_SPHERE;
    else if (ta == JBGPU_SHAPE_PLANE && tb == JBGPU_SHAPE_PLANE) type = JBGPU_PAIR_NONE;
    else if (ta == JBGPU_SHAPE_PLANE || tb == JBGPU_SHAPE_PLANE) type = JBGPU_PAIR_CONVEX_PLANE;
    else type = JBGPU_PAIR_CONVEX_CONVEX;
    const bool planeFirst = type == JBGPU_PAIR_CONVEX_PLANE && ta == JBGPU_SHAPE_PLANE;
    const int b0 = planeFirst ? b : a, b1 = planeFirst ? a : b;
    float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * i;
    if (ldi(m) > 0) {
      if (ldi(m + 1) != b0 || ldi(m + 2) != b1 || ldi(m + 3) != type)
        throw std::invalid_argument("pair " + std::to_string(i) +
                                    ": manifold header is not the dispatcher's (body order or pair type)");
    } else {
      sti(m + 1, b0);
      sti(m + 2, b1);
      sti(m + 3, type);
    }
  }
}

// the pair set (sorted keys with their manifolds) into the context's current pair cache
void uploadPairs(Ctx& c, int64_t n, const uint64_t* keys, const float* manifolds) {
  c.keys[c.cur].ensure(std::max<int64_t>(n, 1));
  c.man[c.cur].ensure((size_t)JBGPU_MANIFOLD_WORDS * std::max<int64_t>(n, 1));
  if (n) {
    JB_CHECK(hipMemcpy(c.keys[c.cur].p, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice));
    JB_CHECK(hipMemcpy(c.man[c.cur].p, manifolds, (size_t)n * JBGPU_MANIFOLD_WORDS * sizeof(float),
                       hipMemcpyHostToDevice));
  }
  c.npairs = n;
  c.npairsOld = 0;
  c.st.pairs = n;
  bpPairTypes(c);
  JB_CHECK(hipStreamSynchronize(c.stream));
}

void stepPhases(Ctx& c, float dt, int last) {
  hipStream_t s = c.stream;
  tBegin(c, T_STEP);
  tBegin(c, T_PREDICT);
  hipLaunchKernelGGL(kPredict, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.grav.p, c.act.p,
                     c.vel.p, (const int*)nullptr);
  tEnd(c, T_PREDICT);
  if (last >= 1) {
    bpUpdateFromBodies(c);
    bpFindPairs(c);
    bpRemapManifolds(c);
  }
  if (last >= 2) npDispatchAll(c);
  if (last >= 3) solverSolve(c, dt);
  if (last >= 4) {
    tBegin(c, T_INTEGRATE);
    if (c.nccd > 0) {  // the motion clamp sweeps against the transforms before integration
      c.trOld.ensure(12 * (size_t)c.nb);
      JB_CHECK(hipMemcpyAsync(c.trOld.p, c.tr.p, 12 * (size_t)c.nb * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(kIntegrate, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.act.p,
                       c.vel.p, (const int*)nullptr);
    ccdClamp(c, dt);
    tEnd(c, T_INTEGRATE);
    hipLaunchKernelGGL(kActivation, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.deactivationDisabled,
                       c.props.p, c.act.p, c.deact.p, c.vel.p, (const int*)nullptr);
  }
  tEnd(c, T_STEP);
  primErrCopy(c);
  JB_CHECK(hipStreamSynchronize(s));
  primErrCheck(c);
  npCollectStats(c);
  solverCollectStats(c);
  if (c.timing) {
    tCollect(c);
    c.timedSteps++;
  }
  if (c.epaOverflow) throw std::length_error("EPA workspace overflow in this step (the step has been applied)");
}

// ---- graph-captured steps ---------------------------------------------------------------------------------
// A step whose every launch is shaped by capacities (proxies, pair capacity) with the per-step counts on the
// device, captured once per (Dbvt stage, pair buffer) parity and replayed: no host wait inside a step and none
// between the steps of one call.  The broadphase runs before the velocity prediction (it reads transforms and
// activation only, the prediction writes velocities only).  A step that finds more pairs than the capacity
// raises a sticky abort flag after its AABB update and changes nothing else; every later replay of the call is
// a no-op, and the host finishes that step on the ordinary path (which raises the capacity) and captures again.
// end of a captured step: the new pair set's count, the step counted, sticky EPA overflow and block-solve errors
__global__ void kGraphStepEnd(int* gc, int nxt, const unsigned long long* __restrict__ npStat,
                              const int* __restrict__ wbInfo) {
  if (!gc[kGcAbort]) {
    gc[kGcPairCount + nxt] = gc[kGcPairs];
    gc[kGcDone]++;
  }
  if (npStat[2]) gc[kGcEpaOverflow] = 1;
  if (wbInfo[0]) gc[kGcWbErr] |= wbInfo[0];
}

std::string graphUnsupported(const Ctx& c) {
  if (c.wbBodies <= 0) return "graph-captured steps need the world-block solve (jbgpu_solver_set_world_blocks)";
  if (c.nj > 0) return "graph-captured steps take no constraints";
  if (!c.deactivationDisabled) return "graph-captured steps need deactivation disabled";
  if (c.tolerance > 0.0f) return "graph-captured steps run a fixed iteration count (no tolerance mode)";
  if (c.useMlp || c.useLearned || c.dumpFeatures) return "graph-captured steps take no learned warm start";
  if (c.timing) return "graph-captured steps are not timed per phase";
  return "";
}

int64_t fbits64(float f) { return (int64_t)__builtin_bit_cast(uint32_t, f); }

// everything a capture bakes into its launches
std::vector<int64_t> graphKey(const Ctx& c, float dt) {
  return {(int64_t)gAllocEpoch, c.stageCurrent, c.cur, c.np, c.nb, c.pairCap, c.nccd, c.numIterations,
          fbits64(c.erp), fbits64(c.warmFactor), fbits64(c.linearSlop), c.warmstart, c.wbBodies, c.wbThreads,
          c.wbPipelined ? 1 : 0, c.wbBlocks, c.wbMaxBodies, c.npairs >= 4 * (int64_t)c.np ? 1 : 0, c.bpMode, c.worldBits, c.maxWorld,
          c.epaSmallFaces, c.epaLargeFaces, c.deactivationDisabled, fbits64(dt)};
}

void stepGraphBody(Ctx& c, float dt) {
  hipStream_t s = c.stream;
  const int* abortFlag = c.gc.p + kGcAbort;
  {  // every counter and flag region the step accumulates into, cleared in one launch
    const int iters = c.numIterations;
    FillList f;
    f.add(c.counter.p, 8, 0u);             // pair counter, traversal stack flag
    f.add(c.counter.p + 4, 4, 0xffffffffu);  // first sorted position per pair type
    f.add(c.counter.p + 6, 2, 0u);          // residual readout
    f.add(c.hugeRange.p, 2 * ((size_t)c.maxWorld + 1), 0u);
    f.add(c.npStat.p, 10, 0u);
    f.add(c.wbInfo.p, 4, 0u);
    f.add(c.resid.p, 2 * (2 * (size_t)iters + 4 + 2 * (size_t)kResidStripes * iters), 0u);
    f.add(c.done.p, 4, 0u);
    graphFillList(s, f);
  }
  const uint64_t* sorted = bpStepGraph(c);
  hipLaunchKernelGGL(kPredict, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.grav.p, c.act.p,
                     c.vel.p, abortFlag);
  bpRemapGraph(c, sorted);
  npDispatchGraph(c);
  solverSolveGraph(c, dt);
  if (c.nccd > 0)
    JB_CHECK(hipMemcpyAsync(c.trOld.p, c.tr.p, 12 * (size_t)c.nb * sizeof(float), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(kIntegrate, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.act.p, c.vel.p,
                     abortFlag);
  ccdClamp(c, dt, abortFlag, c.gc.p + kGcPairs);
  hipLaunchKernelGGL(kActivation, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.deactivationDisabled, c.props.p,
                     c.act.p, c.deact.p, c.vel.p, abortFlag);
  hipLaunchKernelGGL(kGraphStepEnd, dim3(1), dim3(1), 0, s, c.gc.p, c.cur, (const unsigned long long*)c.npStat.p,
                     (const int*)c.wbInfo.p);
}

// the rest of a step whose AABBs are set (a graph-captured step that gave up in its pair search), ordinary path
void stepAfterAabbs(Ctx& c, float dt) {
  hipStream_t s = c.stream;
  hipLaunchKernelGGL(kPredict, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.grav.p, c.act.p,
                     c.vel.p, (const int*)nullptr);
  c.rotationPending = true;
  bpFindPairs(c);
  bpRemapManifolds(c);
  npDispatchAll(c);
  solverSolve(c, dt);
  if (c.nccd > 0) {
    c.trOld.ensure(12 * (size_t)c.nb);
    JB_CHECK(hipMemcpyAsync(c.trOld.p, c.tr.p, 12 * (size_t)c.nb * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  hipLaunchKernelGGL(kIntegrate, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.tr.p, c.props.p, c.act.p, c.vel.p,
                     (const int*)nullptr);
  ccdClamp(c, dt);
  hipLaunchKernelGGL(kActivation, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, dt, c.deactivationDisabled, c.props.p,
                     c.act.p, c.deact.p, c.vel.p, (const int*)nullptr);
  JB_CHECK(hipStreamSynchronize(s));
  npCollectStats(c);
  solverCollectStats(c);
}

void stepGraph(Ctx& c, float dt, int nsteps) {
  const std::string why = graphUnsupported(c);
  if (!why.empty()) throw std::invalid_argument(why);
  hipStream_t s = c.stream;
  bool epaOverflow = false;
  int done = 0;
  while (done < nsteps) {
    if (c.pairCap == 0) {  // no pair search yet: the first step sets the pair count the capacity starts from
      stepPhases(c, dt, 4);
      done++;
      continue;
    }
    // the pair capacity launches are shaped by: 1.5x the last pair count + 4096, rounded up to a sixteenth of a
    // power of two; kept while it stays within [want, 1.25 want], so small changes do not capture again
    const int64_t want = c.npairs + c.npairs / 2 + 4096;
    if (c.pairCap < want || c.pairCap > want + want / 4) {
      int64_t g = 1;
      while (g * 16 < want) g *= 2;
      c.pairCap = (want + g - 1) / g * g;
    }
    if (c.pairCap >= (1ll << 31) / 8) throw std::length_error("pair capacity too large for a graph-captured step");
    c.gc.ensure(kGcWords);
    c.gcHost.ensure(kGcWords);
    bpGraphReserve(c);
    npGraphReserve(c);
    solverGraphReserve(c);
    if (c.nccd > 0) c.trOld.ensure(12 * (size_t)c.nb);
    int* h = c.gcHost.p;
    std::fill(h, h + kGcWords, 0);
    h[kGcPairCount + c.cur] = (int)c.npairs;
    h[kGcPairCount + 1 - c.cur] = (int)c.npairsOld;
    JB_CHECK(hipMemcpyAsync(c.gc.p, h, kGcWords * sizeof(int), hipMemcpyHostToDevice, s));
    JB_CHECK(hipStreamSynchronize(s));  // h is read back into below
    // host state before each replay (stage, pair buffer), and after the last
    std::vector<std::pair<int, int>> state;
    const int todo = nsteps - done;
    for (int k = 0; k < todo; k++) {
      state.emplace_back(c.stageCurrent, c.cur);
      GraphExec& g = c.graphs[2 * (c.stageCurrent & 1) + c.cur];
      const std::vector<int64_t> key = graphKey(c, dt);
      if (!g.e || g.key != key) {
        g.reset();
        const int st = c.stageCurrent, cu = c.cur;
        const bool rp = c.rotationPending;
        const uint64_t epoch = gAllocEpoch;
        hipGraph_t graph = nullptr;
        JB_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          stepGraphBody(c, dt);
        } catch (...) {
          hipGraph_t junk = nullptr;
          (void)hipStreamEndCapture(s, &junk);
          if (junk) (void)hipGraphDestroy(junk);
          c.stageCurrent = st;
          c.cur = cu;
          c.rotationPending = rp;
          throw;
        }
        JB_CHECK(hipStreamEndCapture(s, &graph));
        g.stageAfter = c.stageCurrent;
        g.curAfter = c.cur;
        c.stageCurrent = st;
        c.cur = cu;
        c.rotationPending = rp;
        if (gAllocEpoch != epoch) {
          (void)hipGraphDestroy(graph);
          throw std::logic_error("a device buffer grew while a step was being captured");
        }
        const hipError_t e = hipGraphInstantiate(&g.e, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
          g.e = nullptr;
          throw std::runtime_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        g.key = key;
      }
      JB_CHECK(hipGraphLaunch(g.e, s));
      c.stageCurrent = g.stageAfter;
      c.cur = g.curAfter;
      c.rotationPending = false;
    }
    state.emplace_back(c.stageCurrent, c.cur);
    npStatsCopy(c);
    solverStatsCopy(c);
    primErrCopy(c);
    JB_CHECK(hipMemcpyAsync(h, c.gc.p, kGcWords * sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
    primErrCheck(c);
    const int ok = h[kGcDone];
    epaOverflow = epaOverflow || h[kGcEpaOverflow] != 0;
    if (h[kGcWbErr] & 2) throw std::length_error("a block needed more than 128 colours");
    if (h[kGcWbErr] & 4) throw std::logic_error("a manifold joins two world blocks (bodies of different worlds)");
    if (h[kGcAbort]) {
      // replay `ok` set its AABBs and gave up in its pair search; the replays after it did nothing
      c.stageCurrent = state[ok + 1].first;
      c.cur = state[ok].second;
      c.npairs = h[kGcPairCount + c.cur];
      c.npairsOld = h[kGcPairCount + 1 - c.cur];
      c.npStatPending = false;
      c.solveStat.pending = false;
      c.st.graphRedos++;
      stepAfterAabbs(c, dt);
      epaOverflow = epaOverflow || c.epaOverflow;
      done += ok + 1;
      continue;
    }
    if (ok != todo) throw std::logic_error("a graph-captured step neither completed nor aborted");
    done += ok;
    c.npairs = h[kGcPairCount + c.cur];
    c.npairsOld = h[kGcPairCount + 1 - c.cur];
    c.st.pairs = c.npairs;
    c.nHuge = h[kGcHuge];
    c.st.huge = c.nHuge;
    c.nActive = h[kGcActive];
    c.colorOff.assign({0, (int64_t)c.nActive});
    npCollectStats(c);
    solverCollectStats(c);
    if (c.nActive == 0) {  // as the ordinary step reports a step without a solve
      c.st.iterations = 0;
      c.st.residual = 0.0f;
      c.st.rows = 0;
      c.st.colors = 0;
      c.ncolors = 0;
      c.colorOff.assign(1, 0);
      c.residHist.clear();
    }
  }
  if (epaOverflow) throw std::length_error("EPA workspace overflow in these steps (the steps have been applied)");
}

}  // namespace

extern "C" {

int jbgpu_version(void) { return 1; }

int jbgpu_device_count(int32_t* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (count) *count = n;
  return JBGPU_OK;
}

void jbgpu_config_default(jbgpu_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->bp_mode = JBGPU_BP_DBVT;
  for (int k = 0; k < 3; k++) {
    cfg->world_min[k] = -1000.0f;
    cfg->world_max[k] = 1000.0f;
  }
  cfg->num_iterations = 10;
  cfg->solver_order = JBGPU_ORDER_COLOR;
  cfg->deactivation_disabled = 1;
  cfg->tolerance = 0.0f;
  cfg->max_iterations = 100;
  cfg->warmstart = 1;
  cfg->erp = 0.2f;
  cfg->warmstart_factor = 0.85f;
}

int jbgpu_ctx_create(const jbgpu_config* cfg, jbgpu_ctx** out) {
  if (!out) return JBGPU_ERR_ARG;
  *out = nullptr;
  jbgpu_config d;
  jbgpu_config_default(&d);
  if (!cfg) cfg = &d;
  if (cfg->bp_mode < 0 || cfg->bp_mode > 3 || cfg->num_iterations < 0 || cfg->max_iterations < 1) {
    g_createError = "invalid configuration";
    return JBGPU_ERR_ARG;
  }
  if (cfg->solver_order != JBGPU_ORDER_COLOR) {
    g_createError = "the device solves in graph-colour batches (JBGPU_ORDER_COLOR); the island order with the "
                    "Random(0) shuffle is not available on the device";
    return JBGPU_ERR_ARG;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    g_createError = "no HIP device";
    return JBGPU_ERR_NO_DEVICE;
  }
  jbgpu_ctx* ctx = new jbgpu_ctx();
  Ctx& c = ctx->c;
  c.bpMode = cfg->bp_mode;
  for (int k = 0; k < 3; k++) {
    c.wmin[k] = cfg->world_min[k];
    c.wmax[k] = cfg->world_max[k];
  }
  c.numIterations = cfg->num_iterations;
  c.solverOrder = cfg->solver_order;
  c.deactivationDisabled = cfg->deactivation_disabled;
  c.tolerance = cfg->tolerance;
  c.maxIterations = cfg->max_iterations;
  c.warmstart = cfg->warmstart;
  c.erp = cfg->erp;
  c.warmFactor = cfg->warmstart_factor;
  if (hipGetDevice(&c.device) != hipSuccess || hipStreamCreate(&c.stream) != hipSuccess) {
    g_createError = "hipStreamCreate failed";
    delete ctx;
    return JBGPU_ERR_HIP;
  }
  try {
    mlpStaticInit();  // the device address of the static inference buffer, once per device
  } catch (const std::exception& e) {
    g_createError = e.what();
    (void)hipStreamDestroy(c.stream);
    delete ctx;
    return JBGPU_ERR_HIP;
  }
  bpInitQuant(c);
  *out = ctx;
  return JBGPU_OK;
}

int jbgpu_ctx_destroy(jbgpu_ctx* ctx) {
  if (!ctx) return JBGPU_ERR_ARG;
  Ctx& c = ctx->c;
  (void)hipStreamSynchronize(c.stream);
  // device buffers are freed by their destructors (after the stream has drained)
  for (int r = 0; r < kTimerRegions; r++) {
    if (c.evA[r]) (void)hipEventDestroy(c.evA[r]);
    if (c.evB[r]) (void)hipEventDestroy(c.evB[r]);
  }
  for (hipEvent_t e : c.kEv) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c.stream);
  delete ctx;
  return JBGPU_OK;
}

int jbgpu_last_error(const jbgpu_ctx* ctx, char* buf, int32_t len) {
  if (!buf || len <= 0) return JBGPU_ERR_ARG;
  const std::string& e = ctx ? ctx->c.err : g_createError;
  std::strncpy(buf, e.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return JBGPU_OK;
}

// ---- shape helpers -------------------------------------------------------------
int jbgpu_make_shape(int32_t type, const float* params, const float* verts, float* rec) {
  if (!params || !rec) return JBGPU_ERR_ARG;
  Shape s{};
  s.margin = 0.04f;  // ConvexInternalShape.collisionMargin = CONVEX_DISTANCE_MARGIN
  s.type = type;
  switch (type) {
    case JBGPU_SHAPE_SPHERE:  // SphereShape.java:38-41
      s.dims = v3(params[0], 0, 0);
      s.margin = params[0] * 1.0f;
      break;
    case JBGPU_SHAPE_BOX:  // BoxShape.java:46-50
      s.dims = sub(mulc(v3(params[0], params[1], params[2]), v3(1, 1, 1)), v3(s.margin, s.margin, s.margin));
      break;
    case JBGPU_SHAPE_CAPSULE:  // CapsuleShape.java:54-57
      s.dims = v3(params[0], 0.5f * params[1], params[0]);
      break;
    case JBGPU_SHAPE_HULL: {  // ConvexHullShape.java:45-53 + PolyhedralConvexShape.recalcLocalAabb
      if (!verts) return JBGPU_ERR_ARG;
      s.vfirst = (int)params[0];
      s.vcount = (int)params[1];
      const V3 dirs[6] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
      V3 sup[6];
      float w[6];
      for (int j = 0; j < 6; j++) {
        w[j] = -1e30f;
        sup[j] = v3(0, 0, 0);
      }
      for (int i = 0; i < s.vcount; i++) {
        const float* p = verts + 3 * (s.vfirst + i);
        V3 vtx = v3(p[0] * 1.0f, p[1] * 1.0f, p[2] * 1.0f);
        for (int j = 0; j < 6; j++) {
          float d = dot(dirs[j], vtx);
          if (d > w[j]) {
            sup[j] = vtx;
            w[j] = d;
          }
        }
      }
      s.lmax = v3(sup[0].x + s.margin, sup[1].y + s.margin, sup[2].z + s.margin);
      s.lmin = v3(sup[3].x - s.margin, sup[4].y - s.margin, sup[5].z - s.margin);
      break;
    }
    case JBGPU_SHAPE_PLANE:  // StaticPlaneShape.java:44-47
      s.pn = nor(v3(params[0], params[1], params[2]));
      s.pc = params[3];
      break;
    default:
      return JBGPU_ERR_ARG;
  }
  storeShape(s, rec);
  return JBGPU_OK;
}

int jbgpu_shape_local_inertia(const float* rec, const float* verts, float mass, float* out) {
  if (!rec || !out) return JBGPU_ERR_ARG;
  (void)verts;
  Shape s = loadShape(rec);
  V3 r = v3(0, 0, 0);
  switch (s.type) {
    case JBGPU_SHAPE_BOX: {  // BoxShape.java:154-167
      V3 he = v3(s.dims.x + s.margin, s.dims.y + s.margin, s.dims.z + s.margin);
      float lx = 2.0f * he.x, ly = 2.0f * he.y, lz = 2.0f * he.z;
      r = v3(mass / 12.0f * (ly * ly + lz * lz), mass / 12.0f * (lx * lx + lz * lz), mass / 12.0f * (lx * lx + ly * ly));
      break;
    }
    case JBGPU_SHAPE_SPHERE: {  // SphereShape.java:73-76
      float e = 0.4f * mass * s.margin * s.margin;
      r = v3(e, e, e);
      break;
    }
    case JBGPU_SHAPE_CAPSULE: {  // CapsuleShape.java:126-153
      float rad = s.dims.x;
      V3 he = v3(rad, rad + s.dims.y, rad);
      float m = 0.04f;
      float lx = 2.0f * (he.x + m), ly = 2.0f * (he.y + m), lz = 2.0f * (he.z + m);
      float x2 = lx * lx, y2 = ly * ly, z2 = lz * lz;
      float sm = mass * 0.08333333f;
      r = v3(sm * (y2 + z2), sm * (x2 + z2), sm * (x2 + y2));
      break;
    }
    case JBGPU_SHAPE_HULL: {  // PolyhedralConvexShape.java:134-159 (AABB in the identity frame)
      Tr id;
      id.b = mident();
      id.o = v3(0, 0, 0);
      V3 mn, mx;
      shapeAabb(s, id, mn, mx);
      V3 he = scl(sub(mx, mn), 0.5f);
      float m = s.margin;
      float lx = 2.0f * (he.x + m), ly = 2.0f * (he.y + m), lz = 2.0f * (he.z + m);
      float x2 = lx * lx, y2 = ly * ly, z2 = lz * lz;
      float sm = mass * 0.08333333f;
      r = v3((y2 + z2) * sm, (x2 + z2) * sm, (x2 + y2) * sm);
      break;
    }
    default:
      break;
  }
  st3(out, r);
  return JBGPU_OK;
}

// ---- broadphase ------------------------------------------------------------------
int jbgpu_bp_create_proxies(jbgpu_ctx* ctx, int32_t n, const float* aabb6, const int16_t* group, const int16_t* mask,
                            const int32_t* world, int32_t* uid_out) {
  return guard(ctx, [&] {
    if (n <= 0 || !aabb6) throw std::invalid_argument("bad proxy arguments");
    if (ctx->c.haveWorld) throw std::logic_error("proxies of a world are created by jbgpu_world_create");
    const int n0 = ctx->c.np;
    uploadProxies(ctx->c, n, aabb6, group, mask, world, false);
    if (uid_out)
      for (int i = 0; i < n; i++) uid_out[i] = n0 + i + 1;
  });
}

int jbgpu_bp_set_aabbs(jbgpu_ctx* ctx, int32_t n, const float* aabb6, int32_t on_device) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (n != c.np || !aabb6) throw std::invalid_argument("setAabb needs one box per proxy");
    const float* d = aabb6;
    if (!on_device) {
      c.candC.ensure(6 * (size_t)n);
      JB_CHECK(hipMemcpyAsync(c.candC.p, aabb6, 6 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c.stream));
      d = c.candC.p;
    }
    bpUpdateFromAabbs(c, d);
  });
}

int jbgpu_bp_calculate_overlapping_pairs(jbgpu_ctx* ctx, int64_t* npairs) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (c.np == 0) throw std::logic_error("no proxies");
    tBegin(c, T_STEP);
    bpFindPairs(c);
    bpRemapManifolds(c);
    tEnd(c, T_STEP);
    primErrCopy(c);
    JB_CHECK(hipStreamSynchronize(c.stream));
    primErrCheck(c);
    if (c.timing) {
      tCollect(c);
      c.timedSteps++;
    }
    if (npairs) *npairs = c.npairs;
  });
}

int jbgpu_bp_get_pairs(jbgpu_ctx* ctx, uint64_t* keys, int64_t cap, int64_t* n) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (n) *n = c.npairs;
    if (!keys) return;
    if (cap < c.npairs) throw std::length_error("pair buffer too small");
    if (c.npairs)
      JB_CHECK(hipMemcpy(keys, c.keys[c.cur].p, c.npairs * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

int jbgpu_world_ray_test(jbgpu_ctx* ctx, int32_t n, const float* from3, const float* to3, int16_t group,
                         int16_t mask, const int32_t* world, int32_t* obj, float* fraction, float* normal3,
                         float* point3, int32_t on_device) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (n < 0 || (n > 0 && (!from3 || !to3 || !obj || !fraction || !normal3 || !point3)))
      throw std::invalid_argument("bad ray arguments");
    if (n == 0) return;
    if (world && !on_device)
      for (int i = 0; i < n; i++)
        if (world[i] < 0 || world[i] > c.maxWorld) throw std::invalid_argument("ray world id out of range");
    if (on_device) {
      rayTest(c, n, from3, to3, group, mask, world, obj, fraction, normal3, point3);
      JB_CHECK(hipStreamSynchronize(c.stream));
      return;
    }
    hipStream_t s = c.stream;
    RayTree& t = c.rt;
    t.io.ensure(13 * (size_t)n);
    t.ioInt.ensure(2 * (size_t)n);
    float* dFrom = t.io.p;
    float* dTo = dFrom + 3 * (size_t)n;
    float* dFrac = dTo + 3 * (size_t)n;
    float* dNormal = dFrac + n;
    float* dPoint = dNormal + 3 * (size_t)n;
    int* dObj = t.ioInt.p;
    int* dWorld = world ? t.ioInt.p + n : nullptr;
    JB_CHECK(hipMemcpyAsync(dFrom, from3, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    JB_CHECK(hipMemcpyAsync(dTo, to3, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    if (world) JB_CHECK(hipMemcpyAsync(dWorld, world, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    rayTest(c, n, dFrom, dTo, group, mask, dWorld, dObj, dFrac, dNormal, dPoint);
    JB_CHECK(hipMemcpyAsync(obj, dObj, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(fraction, dFrac, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(normal3, dNormal, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(point3, dPoint, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
  });
}

int jbgpu_bp_destroy_proxies(jbgpu_ctx* ctx, int32_t n, const int32_t* uids) {
  return guard(ctx, [&] {
    if (n < 0 || (n > 0 && !uids)) throw std::invalid_argument("bad proxy arguments");
    if (ctx->c.haveWorld) throw std::logic_error("proxies of a world live as long as the world");
    bpDestroyProxies(ctx->c, n, uids);
  });
}

int jbgpu_bp_get_pair_deltas(jbgpu_ctx* ctx, uint64_t* added, int64_t cap_added, int64_t* n_added, uint64_t* removed,
                             int64_t cap_removed, int64_t* n_removed) {
  return guard(ctx, [&] {
    if (cap_added < 0 || cap_removed < 0) throw std::invalid_argument("negative capacity");
    bpPairDeltas(ctx->c, added, cap_added, n_added, removed, cap_removed, n_removed);
  });
}

int jbgpu_bp_get_proxies(jbgpu_ctx* ctx, float* rec, int32_t* stage_current) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int n = c.np;
    std::vector<int> stage(n), filt(n);
    std::vector<float> leaf(6 * (size_t)n), aabb(6 * (size_t)n);
    JB_CHECK(hipMemcpy(stage.data(), c.stage.p, n * sizeof(int), hipMemcpyDeviceToHost));
    JB_CHECK(hipMemcpy(filt.data(), c.filt.p, n * sizeof(int), hipMemcpyDeviceToHost));
    JB_CHECK(hipMemcpy(leaf.data(), c.leaf.p, leaf.size() * 4, hipMemcpyDeviceToHost));
    JB_CHECK(hipMemcpy(aabb.data(), c.aabb.p, aabb.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
      float* r = rec + (size_t)JBGPU_PROXY_WORDS * i;
      std::memset(r, 0, JBGPU_PROXY_WORDS * 4);
      sti(r, stage[i]);
      for (int k = 0; k < 6; k++) {
        r[1 + k] = leaf[6 * i + k];
        r[7 + k] = aabb[6 * i + k];
      }
      sti(r + 13, (int)(int16_t)(filt[i] & 0xffff));
      sti(r + 14, (int)(int16_t)(filt[i] >> 16));
    }
    if (stage_current) *stage_current = c.stageCurrent;
  });
}

int jbgpu_bp_set_proxies(jbgpu_ctx* ctx, const float* rec, int32_t stage_current) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int n = c.np;
    std::vector<int> stage(n);
    std::vector<float> leaf(6 * (size_t)n), aabb(6 * (size_t)n);
    for (int i = 0; i < n; i++) {
      const float* r = rec + (size_t)JBGPU_PROXY_WORDS * i;
      stage[i] = ldi(r);
      for (int k = 0; k < 6; k++) {
        leaf[6 * i + k] = r[1 + k];
        aabb[6 * i + k] = r[7 + k];
      }
    }
    JB_CHECK(hipMemcpy(c.stage.p, stage.data(), n * sizeof(int), hipMemcpyHostToDevice));
    JB_CHECK(hipMemcpy(c.leaf.p, leaf.data(), leaf.size() * 4, hipMemcpyHostToDevice));
    JB_CHECK(hipMemcpy(c.aabb.p, aabb.data(), aabb.size() * 4, hipMemcpyHostToDevice));
    c.stageCurrent = stage_current;
  });
}

int jbgpu_bp_get_quantized(jbgpu_ctx* ctx, int32_t* q6) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    JB_CHECK(hipMemcpy(q6, c.q.p, 6 * (size_t)c.np * sizeof(int), hipMemcpyDeviceToHost));
  });
}

// ---- world ------------------------------------------------------------------------
int jbgpu_world_create(jbgpu_ctx* ctx, int32_t nshapes, const float* shapes, int32_t nverts, const float* verts,
                       int32_t nbodies, const float* bodies, const int32_t* world_ids) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (nshapes <= 0 || !shapes || nbodies <= 0 || !bodies) throw std::invalid_argument("empty world");
    c.ns = nshapes;
    c.nv = nverts;
    c.nb = nbodies;
    c.nj = 0;  // a new world has no constraints
    c.jointHost.clear();
    c.nccd = 0;  // and no body with continuous collision detection
    c.ccdHost.assign(2 * (size_t)nbodies, 0.0f);
    for (GraphExec& g : c.graphs) g.reset();  // captured steps of an earlier world
    upload(c.shapes, shapes, (size_t)JBGPU_SHAPE_WORDS * nshapes, c.stream);
    std::vector<float> vz(3, 0.0f);
    upload(c.verts, nverts > 0 ? verts : vz.data(), nverts > 0 ? 3 * (size_t)nverts : 3, c.stream);
    c.shapeTypeHost.resize(nshapes);
    for (int k = 0; k < nshapes; k++) c.shapeTypeHost[k] = ldi(shapes + (size_t)JBGPU_SHAPE_WORDS * k);
    bodiesToDevice(c, bodies);
    // addRigidBody -> addCollisionObject: proxy with the shape's AABB (no contact threshold)
    std::vector<float> box(6 * (size_t)nbodies);
    std::vector<int16_t> group(nbodies), mask(nbodies);
    for (int i = 0; i < nbodies; i++) {
      const float* r = bodies + (size_t)JBGPU_BODY_WORDS * i;
      Shape s = loadShape(shapes + JBGPU_SHAPE_WORDS * ldi(r + 27));
      if (s.type == JBGPU_SHAPE_HULL && (s.vfirst < 0 || s.vfirst + s.vcount > nverts))
        throw std::invalid_argument("hull points out of range");
      Tr t;
      for (int k = 0; k < 9; k++) t.b.m[k] = r[k];
      t.o = v3(r[9], r[10], r[11]);
      V3 mn, mx;
      Shape sh = s;
      shapeAabb(sh, t, mn, mx);
      st3(&box[6 * i], mn);
      st3(&box[6 * i + 3], mx);
      bool dyn = (ldi(r + 28) & JBGPU_BODY_STATIC) == 0;
      group[i] = dyn ? 1 : 2;
      mask[i] = dyn ? (int16_t)-1 : (int16_t)(-1 ^ 2);
    }
    // hull AABBs read the hull points through the shape record only (host copy above uses records)
    c.stageCurrent = 0;
    uploadProxies(c, nbodies, box.data(), group.data(), mask.data(), world_ids, true);
    c.worldHost.assign(nbodies, 0);
    if (world_ids) c.worldHost.assign(world_ids, world_ids + nbodies);
    c.wbReady = false;
    c.haveWorld = true;
  });
}

int jbgpu_world_step(jbgpu_ctx* ctx, float dt, int32_t nsteps) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (!(dt > 0.0f) || nsteps < 0) throw std::invalid_argument("bad step arguments");
    if (c.stepGraph) {
      stepGraph(c, dt, nsteps);
      return;
    }
    for (int i = 0; i < nsteps; i++) stepPhases(c, dt, 4);
  });
}

int jbgpu_world_set_step_graph(jbgpu_ctx* ctx, int32_t enable) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (enable) {
      requireWorld(c);
      const std::string why = graphUnsupported(c);
      if (!why.empty()) throw std::invalid_argument(why);
    }
    c.stepGraph = enable != 0;
    if (!enable)
      for (GraphExec& g : c.graphs) g.reset();
  });
}

int jbgpu_world_step_until(jbgpu_ctx* ctx, float dt, int32_t phase) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (phase < 0 || phase > 4) throw std::invalid_argument("phase must be 0..4");
    stepPhases(c, dt, phase);
  });
}

}  // extern "C"

namespace {
// bodies of n new constraints: A in range, B in range or -1 (the fixed body), A != B; all checked before any change
void checkJointBodies(Ctx& c, int n, const int32_t* body_a, const int32_t* body_b) {
  if ((int64_t)c.nj + n > (1 << 30)) throw std::length_error("too many constraints");
  for (int i = 0; i < n; i++) {
    const int a = body_a[i], b = body_b[i];
    if (a < 0 || a >= c.nb) throw std::invalid_argument("constraint " + std::to_string(i) + ": body A out of range");
    if (b < -1 || b >= c.nb)
      throw std::invalid_argument("constraint " + std::to_string(i) + ": body B out of range (-1: the fixed body)");
    if (a == b) throw std::invalid_argument("constraint " + std::to_string(i) + ": body A is body B");
  }
}
// the last n host records to the device, appliedImpulse 0 (TypedConstraint's initial value)
void appendJoints(Ctx& c, int n) {
  if (n == 0) return;
  const int nj = c.nj + n;
  c.joints.ensureKeep((size_t)kJointWords * nj);
  c.jointState.ensureKeep((size_t)kJointStateWords * nj);
  JB_CHECK(hipMemcpy(c.joints.p + (size_t)kJointWords * c.nj, c.jointHost.data() + (size_t)kJointWords * c.nj,
                     (size_t)kJointWords * n * sizeof(float), hipMemcpyHostToDevice));
  JB_CHECK(hipMemset(c.jointState.p + (size_t)kJointStateWords * c.nj, 0, (size_t)kJointStateWords * n * sizeof(float)));
  c.nj = nj;
  jointsChanged(c);
}
Tr trFrom12(const float* p) {
  Tr t;
  for (int k = 0; k < 9; k++) t.b.m[k] = p[k];
  t.o = v3(p[9], p[10], p[11]);
  return t;
}
void trTo12(const Tr& t, float* p) {
  for (int k = 0; k < 9; k++) p[k] = t.b.m[k];
  p[9] = t.o.x;
  p[10] = t.o.y;
  p[11] = t.o.z;
}
V3 columnOf(const M3& m, int c) { return v3(m.m[c], m.m[3 + c], m.m[6 + c]); }  // MatrixUtil.getColumn
void setRowOf(M3& m, int r, float x, float y, float z) {                         // MatrixUtil.setRow
  m.m[3 * r] = x;
  m.m[3 * r + 1] = y;
  m.m[3 * r + 2] = z;
}
// HingeConstraint's constructors (HingeConstraint.java:84-238), host side: the frames the solver uses
void hingeFramesFromAxes(const Tr& trA, V3 pivotA, V3 pivotB, V3 axisA, V3 axisB, bool fixedB, Tr& frameA,
                         Tr& frameB) {
  const float eps = 1.19209290e-07f;  // BulletGlobals.SIMD_EPSILON
  V3 a1 = columnOf(trA.b, 0), a2;
  if (!fixedB) {
    const float projection = dot(axisA, a1);
    if (projection >= 1.0f - eps) {
      a1 = scl(columnOf(trA.b, 2), -1.0f);
      a2 = columnOf(trA.b, 1);
    } else if (projection <= -1.0f + eps) {
      a1 = columnOf(trA.b, 2);
      a2 = columnOf(trA.b, 1);
    } else {  // not renormalised: both axes have length sin(angle between axisA and A's x axis)
      a2 = crs(axisA, a1);
      a1 = crs(a2, axisA);
    }
  } else {
    const float projection = dot(a1, axisA);
    if (projection > eps) {
      a1 = scl(a1, projection);
      a1 = sub(a1, axisA);
    } else {
      a1 = columnOf(trA.b, 1);
    }
    a2 = crs(axisA, a1);
    axisB = mulMV(trA.b, scl(axisA, -1.0f));  // the axis in the fixed body: -axisA in world space
  }
  frameA.o = pivotA;
  setRowOf(frameA.b, 0, a1.x, a2.x, axisA.x);
  setRowOf(frameA.b, 1, a1.y, a2.y, axisA.y);
  setRowOf(frameA.b, 2, a1.z, a2.z, axisA.z);
  // QuaternionUtil.shortestArcQuat (QuaternionUtil.java:49-68; its antiparallel test is a double comparison),
  // then quatRotate (:78-90): (q * v) * conj(q)
  const V3 cr = crs(axisA, axisB);
  const float d = dot(axisA, axisB);
  Q4 arc;
  if ((double)d < -1.0 + (double)eps) {
    arc = Q4{0.0f, 1.0f, 0.0f, 0.0f};
  } else {
    const float sq = (float)std::sqrt((double)((1.0f + d) * 2.0f));
    const float rs = 1.0f / sq;
    arc = Q4{cr.x * rs, cr.y * rs, cr.z * rs, sq * 0.5f};
  }
  Q4 q{arc.w * a1.x + arc.y * a1.z - arc.z * a1.y, arc.w * a1.y + arc.z * a1.x - arc.x * a1.z,
       arc.w * a1.z + arc.x * a1.y - arc.y * a1.x, -arc.x * a1.x - arc.y * a1.y - arc.z * a1.z};
  q = qmul(q, Q4{-arc.x, -arc.y, -arc.z, arc.w});
  const V3 b1 = v3(q.x, q.y, q.z);
  const V3 b2 = crs(axisB, b1);
  // the two-body form flips B's axis column; the fixed-body form keeps it (it is -axisA already)
  frameB.o = fixedB ? xform(trA, pivotA) : pivotB;
  setRowOf(frameB.b, 0, b1.x, b2.x, fixedB ? axisB.x : -axisB.x);
  setRowOf(frameB.b, 1, b1.y, b2.y, fixedB ? axisB.y : -axisB.y);
  setRowOf(frameB.b, 2, b1.z, b2.z, fixedB ? axisB.z : -axisB.z);
}
void hingeFramesFromFrames(const Tr& trA, const Tr& inA, const Tr& inB, bool fixedB, Tr& frameA, Tr& frameB) {
  frameA = inA;
  frameB = fixedB ? inA : inB;
  frameB.b.m[2] *= -1.0f;  // flip axis: M02, M12, M22
  frameB.b.m[5] *= -1.0f;
  frameB.b.m[8] *= -1.0f;
  if (fixedB) frameB.o = xform(trA, inA.o);
}
}  // namespace

extern "C" {

int jbgpu_world_add_point2point(jbgpu_ctx* ctx, int32_t n, const int32_t* body_a, const int32_t* body_b,
                                const float* pivot_a, const float* pivot_b, const float* settings) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (n < 0 || (n > 0 && (!body_a || !body_b || !pivot_a || !pivot_b)))
      throw std::invalid_argument("bad constraint arguments");
    if ((int64_t)c.nj + n > (1 << 30)) throw std::length_error("too many constraints");
    checkJointBodies(c, n, body_a, body_b);
    for (int i = 0; i < n; i++) {
      float r[kJointWords] = {};
      sti(r, kJointP2P);
      sti(r + 1, body_a[i]);
      sti(r + 2, body_b[i]);
      for (int k = 0; k < 3; k++) {
        r[4 + k] = pivot_a[3 * i + k];
        r[7 + k] = pivot_b[3 * i + k];
      }
      r[10] = settings ? settings[3 * i] : 0.3f;  // ConstraintSetting defaults (Point2PointConstraint.java:202-206)
      r[11] = settings ? settings[3 * i + 1] : 1.0f;
      r[12] = settings ? settings[3 * i + 2] : 0.0f;
      c.jointHost.insert(c.jointHost.end(), r, r + kJointWords);
    }
    appendJoints(c, n);
  });
}

int jbgpu_world_add_hinge(jbgpu_ctx* ctx, int32_t n, const int32_t* body_a, const int32_t* body_b,
                          const float* frame_a, const float* frame_b, const float* params) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (n < 0 || (n > 0 && (!body_a || !body_b || !frame_a || !frame_b)))
      throw std::invalid_argument("bad constraint arguments");
    checkJointBodies(c, n, body_a, body_b);
    for (int i = 0; i < n; i++) {
      float r[kJointWords] = {};
      sti(r, kJointHinge);
      sti(r + 1, body_a[i]);
      sti(r + 2, body_b[i]);
      for (int k = 0; k < 12; k++) {
        r[4 + k] = frame_a[12 * i + k];
        r[16 + k] = frame_b[12 * i + k];
      }
      // free hinge defaults (HingeConstraint.java:123-129): no limit, softness 0.9, bias 0.3, relaxation 1
      const float def[9] = {1e30f, -1e30f, 0.9f, 0.3f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      for (int k = 0; k < 9; k++) r[28 + k] = params ? params[9 * i + k] : def[k];
      r[33] = r[33] != 0.0f ? 1.0f : 0.0f;
      r[34] = r[34] != 0.0f ? 1.0f : 0.0f;
      c.jointHost.insert(c.jointHost.end(), r, r + kJointWords);
    }
    appendJoints(c, n);
  });
}

int jbgpu_hinge_frames(int32_t from_frames, int32_t fixed_b, const float* tr_a, const float* pivot_a,
                       const float* pivot_b, const float* axis_a, const float* axis_b, const float* frame_a_in,
                       const float* frame_b_in, float* frame_a, float* frame_b) {
  try {
    if (!tr_a || !frame_a || !frame_b) throw std::invalid_argument("null argument");
    Tr ta = trFrom12(tr_a), fa, fb;
    if (from_frames) {
      if (!frame_a_in || (!fixed_b && !frame_b_in)) throw std::invalid_argument("null frame");
      const Tr ia = trFrom12(frame_a_in);
      hingeFramesFromFrames(ta, ia, fixed_b ? ia : trFrom12(frame_b_in), fixed_b != 0, fa, fb);
    } else {
      if (!pivot_a || !axis_a || (!fixed_b && (!pivot_b || !axis_b))) throw std::invalid_argument("null pivot or axis");
      const V3 z = v3(0.0f, 0.0f, 0.0f);
      hingeFramesFromAxes(ta, ld3(pivot_a), fixed_b ? z : ld3(pivot_b), ld3(axis_a), fixed_b ? z : ld3(axis_b),
                          fixed_b != 0, fa, fb);
    }
    trTo12(fa, frame_a);
    trTo12(fb, frame_b);
    return JBGPU_OK;
  } catch (const std::exception& e) {
    g_createError = e.what();
    return JBGPU_ERR_ARG;
  }
}

int jbgpu_np_set_epa_capacity(jbgpu_ctx* ctx, int32_t small_faces, int32_t large_faces) {
  return guard(ctx, [&] {
    if (small_faces < 16 || small_faces > kEpaSmallFaces || large_faces < 16 || large_faces > kEpaFaces)
      throw std::invalid_argument("EPA face capacities must lie in [16, " + std::to_string(kEpaSmallFaces) + "] and [16, " +
                                  std::to_string(kEpaFaces) + "]");
    ctx->c.epaSmallFaces = small_faces;
    ctx->c.epaLargeFaces = large_faces;
  });
}

int jbgpu_prim_sort(jbgpu_ctx* ctx, void* keys, int32_t key_bytes, int32_t* values, int64_t n, int32_t bit_lo,
                    int32_t bit_hi) {
  return guard(ctx, [&] {
    if (n < 0 || n > ((int64_t)1 << 28)) throw std::invalid_argument("sort length must lie in [0, 2^28]");
    if (key_bytes != 4 && key_bytes != 8) throw std::invalid_argument("key size must be 4 or 8 bytes");
    if (bit_lo < 0 || bit_hi > 8 * key_bytes || bit_lo > bit_hi)
      throw std::invalid_argument("bit range must satisfy 0 <= bit_lo <= bit_hi <= 8 * key_bytes");
    if (n > 0 && !keys) throw std::invalid_argument("null keys");
    if (n == 0) return;
    Ctx& c = ctx->c;
    hipStream_t s = c.stream;
    auto run = [&](auto* hk) {
      using K = std::remove_pointer_t<decltype(hk)>;
      DBuf<K> k0, k1;
      DBuf<int> v0, v1;
      k0.ensure(n);
      k1.ensure(n);
      JB_CHECK(hipMemcpyAsync(k0.p, hk, n * sizeof(K), hipMemcpyHostToDevice, s));
      Db<K> dk(k0.p, k1.p);
      if (values) {
        v0.ensure(n);
        v1.ensure(n);
        JB_CHECK(hipMemcpyAsync(v0.p, values, n * sizeof(int), hipMemcpyHostToDevice, s));
        Db<int> dv(v0.p, v1.p);
        primSortPairs(c, dk, dv, (int)n, bit_lo, bit_hi);
        JB_CHECK(hipMemcpyAsync(values, dv.Current(), n * sizeof(int), hipMemcpyDeviceToHost, s));
      } else {
        primSortKeys(c, dk, (int)n, bit_lo, bit_hi);
      }
      JB_CHECK(hipMemcpyAsync(hk, dk.Current(), n * sizeof(K), hipMemcpyDeviceToHost, s));
      primErrCopy(c);
      JB_CHECK(hipStreamSynchronize(s));
      primErrCheck(c);
    };
    if (key_bytes == 8)
      run((uint64_t*)keys);
    else
      run((int*)keys);
  });
}

int jbgpu_world_num_joints(jbgpu_ctx* ctx, int32_t* n) {
  return guard(ctx, [&] {
    if (n) *n = ctx->c.nj;
  });
}

int jbgpu_world_joint_impulses(jbgpu_ctx* ctx, float* out) {
  return guard(ctx, [&] {
    if (ctx->c.nj > 0 && !out) throw std::invalid_argument("null output");
    jointsImpulses(ctx->c, out);
  });
}

int jbgpu_world_joint_order(jbgpu_ctx* ctx, int32_t* order, int32_t cap, int32_t* n) {
  return guard(ctx, [&] {
    if (cap < 0) throw std::invalid_argument("negative capacity");
    const int m = jointsOrderSolved(ctx->c, order, cap);
    if (n) *n = m;
  });
}

int jbgpu_world_set_ccd(jbgpu_ctx* ctx, int32_t n, const int32_t* bodies, const float* motion_threshold,
                        const float* swept_sphere_radius) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (n < 0 || (n > 0 && (!bodies || !motion_threshold || !swept_sphere_radius)))
      throw std::invalid_argument("bad CCD arguments");
    for (int k = 0; k < n; k++) {
      if (bodies[k] < 0 || bodies[k] >= c.nb) throw std::invalid_argument("CCD body index out of range");
      if (!std::isfinite(motion_threshold[k]) || !std::isfinite(swept_sphere_radius[k]) ||
          swept_sphere_radius[k] < 0.0f)
        throw std::invalid_argument("CCD motion threshold and swept sphere radius must be finite, the radius >= 0");
    }
    for (int k = 0; k < n; k++) {
      c.ccdHost[2 * (size_t)bodies[k]] = motion_threshold[k];
      c.ccdHost[2 * (size_t)bodies[k] + 1] = swept_sphere_radius[k];
    }
    std::vector<int> list;
    for (int i = 0; i < c.nb; i++)
      if (c.ccdHost[2 * (size_t)i] != 0.0f) list.push_back(i);
    c.nccd = (int)list.size();
    upload(c.ccd, c.ccdHost.data(), c.ccdHost.size(), c.stream);
    upload(c.ccdList, list.data(), list.size(), c.stream);
    JB_CHECK(hipStreamSynchronize(c.stream));
  });
}

int jbgpu_world_convex_sweep_test(jbgpu_ctx* ctx, const float* cast_shape, int32_t n, const float* from12,
                                  const float* to12, int16_t group, int16_t mask, const int32_t* world, int32_t* obj,
                                  float* fraction, float* normal3, float* point3, int32_t on_device) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (!cast_shape || n < 0 || (n > 0 && (!from12 || !to12 || !obj || !fraction || !normal3 || !point3)))
      throw std::invalid_argument("bad sweep arguments");
    const Shape cs = loadShape(cast_shape);
    if (cs.type != JBGPU_SHAPE_SPHERE && cs.type != JBGPU_SHAPE_BOX && cs.type != JBGPU_SHAPE_CAPSULE &&
        cs.type != JBGPU_SHAPE_HULL)
      throw std::invalid_argument("the cast shape must be convex (sphere, box, capsule or hull)");
    if (cs.type == JBGPU_SHAPE_HULL && (cs.vfirst < 0 || cs.vcount <= 0 || cs.vfirst + cs.vcount > c.nv))
      throw std::invalid_argument("cast hull points out of the world's point table");
    if (n == 0) return;
    if (world && !on_device)
      for (int i = 0; i < n; i++)
        if (world[i] < 0 || world[i] > c.maxWorld) throw std::invalid_argument("sweep world id out of range");
    if (on_device) {
      sweepTest(c, n, cast_shape, from12, to12, group, mask, world, obj, fraction, normal3, point3);
      JB_CHECK(hipStreamSynchronize(c.stream));
      return;
    }
    hipStream_t s = c.stream;
    RayTree& t = c.rt;
    t.io.ensure(31 * (size_t)n);
    t.ioInt.ensure(2 * (size_t)n);
    float* dFrom = t.io.p;
    float* dTo = dFrom + 12 * (size_t)n;
    float* dFrac = dTo + 12 * (size_t)n;
    float* dNormal = dFrac + n;
    float* dPoint = dNormal + 3 * (size_t)n;
    int* dObj = t.ioInt.p;
    int* dWorld = world ? t.ioInt.p + n : nullptr;
    JB_CHECK(hipMemcpyAsync(dFrom, from12, 12 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    JB_CHECK(hipMemcpyAsync(dTo, to12, 12 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    if (world) JB_CHECK(hipMemcpyAsync(dWorld, world, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    sweepTest(c, n, cast_shape, dFrom, dTo, group, mask, dWorld, dObj, dFrac, dNormal, dPoint);
    JB_CHECK(hipMemcpyAsync(obj, dObj, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(fraction, dFrac, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(normal3, dNormal, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(point3, dPoint, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
  });
}

int jbgpu_world_get_bodies(jbgpu_ctx* ctx, float* bodies) {
  return guard(ctx, [&] {
    requireWorld(ctx->c);
    bodiesFromDevice(ctx->c, bodies);
  });
}

int jbgpu_world_set_bodies(jbgpu_ctx* ctx, const float* bodies) {
  return guard(ctx, [&] {
    requireWorld(ctx->c);
    bodiesToDevice(ctx->c, bodies);
  });
}

int jbgpu_world_num_pairs(jbgpu_ctx* ctx, int64_t* n) {
  return guard(ctx, [&] {
    if (n) *n = ctx->c.npairs;
  });
}

int jbgpu_world_get_pairs(jbgpu_ctx* ctx, uint64_t* keys, float* manifolds) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (c.npairs == 0) return;
    if (keys) JB_CHECK(hipMemcpy(keys, c.keys[c.cur].p, c.npairs * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (manifolds)
      JB_CHECK(hipMemcpy(manifolds, c.man[c.cur].p, (size_t)c.npairs * JBGPU_MANIFOLD_WORDS * sizeof(float),
                         hipMemcpyDeviceToHost));
  });
}

int jbgpu_world_set_pairs(jbgpu_ctx* ctx, int64_t n, const uint64_t* keys, const float* manifolds) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    checkPairRecords(c.nb, n, keys, manifolds);
    uploadPairs(c, n, keys, manifolds);
  });
}

// ---- Dispatcher seam ---------------------------------------------------------------
int jbgpu_np_dispatch(jbgpu_ctx* ctx, int64_t n, const uint64_t* keys, const float* manifolds_in,
                      float* manifolds_out) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (n > 0 && !manifolds_out) throw std::invalid_argument("null output manifolds");
    checkPairRecords(c.nb, n, keys, manifolds_in);
    std::vector<float> man(manifolds_in, manifolds_in + (size_t)JBGPU_MANIFOLD_WORDS * (n > 0 ? n : 0));
    canonicalHeaders(c, n, keys, man.data());
    uploadPairs(c, n, keys, man.data());
    npDispatchAll(c);
    JB_CHECK(hipStreamSynchronize(c.stream));
    npCollectStats(c);
    if (n)
      JB_CHECK(hipMemcpy(manifolds_out, c.man[c.cur].p, (size_t)n * JBGPU_MANIFOLD_WORDS * sizeof(float),
                         hipMemcpyDeviceToHost));
    if (c.epaOverflow)
      throw std::length_error("EPA workspace overflow in this pass (the manifolds have been updated)");
  });
}

// ---- ConstraintSolver seam -----------------------------------------------------------
int jbgpu_solver_prepare(jbgpu_ctx* ctx, int32_t num_bodies, int64_t num_manifolds) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    requireWorld(c);
    if (num_bodies != c.nb) throw std::invalid_argument("prepareSolve: body count differs from the world's");
    if (num_manifolds < 0 || num_manifolds != c.npairs)
      throw std::invalid_argument("prepareSolve: manifold count differs from the pair cache's");
    c.islandManifolds.clear();
    c.solveBatchOpen = true;
  });
}

int jbgpu_solver_add_island(jbgpu_ctx* ctx, int64_t n, const int32_t* pair_index) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (!c.solveBatchOpen) throw std::logic_error("solveGroup outside prepareSolve / allSolved");
    if (n < 0 || (n > 0 && !pair_index)) throw std::invalid_argument("bad island arguments");
    for (int64_t i = 0; i < n; i++)
      if (pair_index[i] < 0 || pair_index[i] >= c.npairs) throw std::invalid_argument("manifold index out of range");
    c.islandManifolds.insert(c.islandManifolds.end(), pair_index, pair_index + n);
  });
}

int jbgpu_solver_all_solved(jbgpu_ctx* ctx, float dt) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (!c.solveBatchOpen) throw std::logic_error("allSolved without prepareSolve");
    if (!(dt > 0.0f)) throw std::invalid_argument("bad time step");
    c.solveBatchOpen = false;
    std::vector<char> seen(c.npairs, 0);
    for (int k : c.islandManifolds) {
      if (seen[k]) throw std::invalid_argument("a manifold was recorded in two islands");
      seen[k] = 1;
    }
    solverSolve(c, dt, &c.islandManifolds);
    JB_CHECK(hipStreamSynchronize(c.stream));
    solverCollectStats(c);
  });
}

int jbgpu_check_pair_records(int32_t nbodies, int64_t n, const uint64_t* keys, const float* manifolds) {
  try {
    if (nbodies < 0) throw std::invalid_argument("negative body count");
    checkPairRecords(nbodies, n, keys, manifolds);
    return JBGPU_OK;
  } catch (const std::exception& e) {
    g_createError = e.what();
    return JBGPU_ERR_ARG;
  }
}

int jbgpu_world_stats(jbgpu_ctx* ctx, jbgpu_stats* out) {
  return guard(ctx, [&] {
    const Stats& s = ctx->c.st;
    out->pairs = s.pairs;
    out->contacts = ctx->c.nb > 0 ? npCountContacts(ctx->c) : 0;
    out->manifolds = ctx->c.nActive;
    out->colors = s.colors;
    out->epa_calls = s.epaCalls;
    out->epa_fails = s.epaFails;
    out->iterations = s.iterations;
    out->residual = s.residual;
    out->huge_proxies = s.huge;
    out->rows = s.rows;
    out->epa_retries = s.epaRetries;
    out->epa_huge = s.epaHuge;
    out->graph_redos = s.graphRedos;
  });
}

// ---- solver ------------------------------------------------------------------------
int jbgpu_solver_get_order(jbgpu_ctx* ctx, int32_t* order, int64_t cap, int64_t* n, int64_t* offsets,
                           int32_t* ncolors) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int64_t na = c.nActive;
    if (n) *n = na;
    if (ncolors) *ncolors = c.ncolors;
    if (offsets)
      for (int k = 0; k <= c.ncolors; k++) offsets[k] = c.colorOff[k];
    if (!order || na == 0) return;
    if (cap < na) throw std::length_error("order buffer too small");
    std::vector<int> co(na), act(na);
    JB_CHECK(hipMemcpy(co.data(), c.colorOrder.p, na * sizeof(int), hipMemcpyDeviceToHost));
    JB_CHECK(hipMemcpy(act.data(), c.mActive.p, na * sizeof(int), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < na; k++) order[k] = act[co[k]];
  });
}

int jbgpu_solver_set_warmstart(jbgpu_ctx* ctx, const float* init, int64_t nslots, int32_t on_device) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (!init) {
      c.useLearned = false;
      return;
    }
    c.learned.ensure(3 * (size_t)std::max<int64_t>(nslots, 1));
    JB_CHECK(hipMemcpy(c.learned.p, init, 3 * (size_t)nslots * sizeof(float),
                       on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    c.useLearned = true;
    c.learnedSlots = nslots;
  });
}

int jbgpu_solver_enable_features(jbgpu_ctx* ctx, int32_t on) {
  return guard(ctx, [&] { ctx->c.dumpFeatures = on != 0; });
}

int jbgpu_solver_get_features(jbgpu_ctx* ctx, float* out, int64_t nslots) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (!c.dumpFeatures) throw std::logic_error("features not enabled");
    if (nslots > 4 * c.nActive) throw std::length_error("more slots than the last solve had");
    if (nslots) JB_CHECK(hipMemcpy(out, c.features.p, 16 * (size_t)nslots * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int jbgpu_solver_get_impulses(jbgpu_ctx* ctx, float* out, int64_t nslots) {
  return guard(ctx, [&] {
    if (nslots > 0 && !out) throw std::invalid_argument("null output");
    solverGetImpulses(ctx->c, out, nslots);
  });
}

int jbgpu_solver_set_params(jbgpu_ctx* ctx, float tolerance, int32_t max_iterations, int32_t num_iterations) {
  return guard(ctx, [&] {
    if (max_iterations < 1 || num_iterations < 0) throw std::invalid_argument("bad iteration counts");
    ctx->c.tolerance = tolerance;
    ctx->c.maxIterations = max_iterations;
    ctx->c.numIterations = num_iterations;
  });
}

int jbgpu_solver_set_blocks(jbgpu_ctx* ctx, int32_t block_size) {
  return guard(ctx, [&] {
    if (block_size < 0 || block_size > 2048) throw std::invalid_argument("block size must be in [0, 2048]");
    ctx->c.blockSize = block_size;
  });
}

int jbgpu_solver_set_world_blocks(jbgpu_ctx* ctx, int32_t bodies_per_block, int32_t threads, int32_t pipelined) {
  return guard(ctx, [&] {
    if (bodies_per_block != 0 && (bodies_per_block < 16 || bodies_per_block > 1024))
      throw std::invalid_argument("bodies per block must be 0 or in [16, 1024]");
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
      throw std::invalid_argument("threads per block must be 64, 128, 256 or 512");
    Ctx& c = ctx->c;
    c.wbBodies = bodies_per_block;
    c.wbThreads = threads;
    c.wbPipelined = pipelined != 0;
    c.wbReady = false;
  });
}

int jbgpu_solver_set_colouring(jbgpu_ctx* ctx, int32_t degree_first, int32_t recolour_passes) {
  return guard(ctx, [&] {
    if (recolour_passes < 0 || recolour_passes > 8) throw std::invalid_argument("recolour passes must be in [0, 8]");
    ctx->c.colourByDegree = degree_first != 0;
    ctx->c.recolourPasses = recolour_passes;
  });
}

int jbgpu_solver_set_warmstart_factor(jbgpu_ctx* ctx, float factor) {
  return guard(ctx, [&] {
    if (!std::isfinite(factor)) throw std::invalid_argument("warm-start factor must be finite");
    ctx->c.warmFactor = factor;
  });
}

int jbgpu_solver_set_tail(jbgpu_ctx* ctx, int32_t threshold) {
  return guard(ctx, [&] {
    if (threshold < 0) throw std::invalid_argument("negative tail threshold");
    ctx->c.tailThreshold = threshold;
  });
}

int jbgpu_solver_residual_history(jbgpu_ctx* ctx, float* out, int32_t cap, int32_t* n) {
  return guard(ctx, [&] {
    const auto& h = ctx->c.residHist;
    if (n) *n = (int32_t)h.size();
    for (int32_t k = 0; k < cap && k < (int32_t)h.size(); k++) out[k] = h[k];
  });
}

// ---- learned warm start ----------------------------------------------------------------
static_assert(jbg::kTimerRegions == JBGPU_TIMER_REGIONS, "timer regions");
static_assert(jbg::kMlpParams == JBGPU_MLP_PARAMS && jbg::kMlpNormWords == JBGPU_MLP_NORM_WORDS, "MLP layout");
static_assert(jbg::kStaticInferMax == JBGPU_MLP_STATIC_MAX, "static inference limit");

int jbgpu_mlp_set_params(jbgpu_ctx* ctx, const float* params, const float* norm) {
  return guard(ctx, [&] { mlpSetParams(ctx->c, params, norm); });
}

int jbgpu_mlp_get_params(jbgpu_ctx* ctx, float* params, float* norm) {
  return guard(ctx, [&] { mlpGetParams(ctx->c, params, norm); });
}

int jbgpu_mlp_infer(jbgpu_ctx* ctx, const float* x, int64_t n, float* y) {
  return guard(ctx, [&] { mlpInfer(ctx->c, x, n, y); });
}

int jbgpu_mlp_infer_static(jbgpu_ctx* ctx, const float* x, int64_t n, float* y) {
  return guard(ctx, [&] { mlpInferStatic(ctx->c, x, n, y); });
}

int jbgpu_mlp_grad(jbgpu_ctx* ctx, const float* x, const float* t, int64_t n, float* grad, float* loss) {
  return guard(ctx, [&] { mlpGrad(ctx->c, x, t, n, grad, loss); });
}

int jbgpu_mlp_train(jbgpu_ctx* ctx, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float* loss_per_epoch) {
  return guard(ctx, [&] { mlpTrain(ctx->c, x, t, n, epochs, batch, lr, seed, loss_per_epoch); });
}

int jbgpu_solver_use_mlp(jbgpu_ctx* ctx, int32_t on) {
  return guard(ctx, [&] {
    if (on < 0 || on > 2) throw std::invalid_argument("mode must be 0, 1 or 2");
    if (on && !ctx->c.mlpReady) throw std::logic_error("no model loaded or trained");
    ctx->c.useMlp = on != 0;
    ctx->c.mlpNewOnly = on == 2;
  });
}

// ---- timing -------------------------------------------------------------------------
int jbgpu_timing_enable(jbgpu_ctx* ctx, int32_t on) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (on && !c.evA[0])
      for (int r = 0; r < kTimerRegions; r++) {
        JB_CHECK(hipEventCreate(&c.evA[r]));
        JB_CHECK(hipEventCreate(&c.evB[r]));
      }
    c.timing = on != 0;
    for (int r = 0; r < kTimerRegions; r++) {
      c.ms[r] = 0;
      c.launches[r] = 0;
      c.pending[r] = false;
    }
    c.timedSteps = 0;
  });
}

int jbgpu_buffer_alloc(jbgpu_ctx* ctx, int64_t bytes, void** dev) {
  return guard(ctx, [&] {
    if (bytes <= 0 || !dev) throw std::invalid_argument("bad buffer size");
    *dev = nullptr;
    JB_CHECK(hipMalloc(dev, (size_t)bytes));
  });
}

int jbgpu_buffer_upload(jbgpu_ctx* ctx, void* dev, const void* host, int64_t bytes) {
  return guard(ctx, [&] {
    if (!dev || !host || bytes < 0) throw std::invalid_argument("bad upload");
    JB_CHECK(hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->c.stream));
    JB_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int jbgpu_buffer_free(jbgpu_ctx* ctx, void* dev) {
  return guard(ctx, [&] {
    JB_CHECK(hipStreamSynchronize(ctx->c.stream));
    if (dev) JB_CHECK(hipFree(dev));
  });
}

int jbgpu_measure_copy(jbgpu_ctx* ctx, int64_t bytes, int32_t reps, double* gbs) {
  return guard(ctx, [&] {
    if (bytes < 16 || reps < 1 || !gbs) throw std::invalid_argument("bad copy probe arguments");
    Ctx& c = ctx->c;
    const int64_t n4 = bytes / 16;
    DBuf<float4> a, b;
    a.ensure(n4);
    b.ensure(n4);
    JB_CHECK(hipMemsetAsync(a.p, 0, n4 * 16, c.stream));
    hipEvent_t e0, e1;
    JB_CHECK(hipEventCreate(&e0));
    JB_CHECK(hipEventCreate(&e1));
    const int grid = 8 * 256 * 4;  // four workgroups per SIMD-pair across the 256 CUs, grid-stride
    hipLaunchKernelGGL(kCopyProbe, dim3(grid), dim3(256), 0, c.stream, n4, (const float4*)a.p, b.p);  // warm
    JB_CHECK(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(kCopyProbe, dim3(grid), dim3(256), 0, c.stream, n4, (const float4*)a.p, b.p);
    JB_CHECK(hipEventRecord(e1, c.stream));
    JB_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    JB_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = 2.0 * 16.0 * (double)n4 * reps / (ms * 1e-3) / 1e9;  // bytes read + written
  });
}

int jbgpu_timing_kernels(jbgpu_ctx* ctx, int32_t every) {
  return guard(ctx, [&] {
    if (every < 0) throw std::invalid_argument("negative sampling period");
    Ctx& c = ctx->c;
    c.kernelEvery = every;
    c.kernelSteps = 0;
    for (int k = 0; k < 4; k++) {
      c.kMs[k] = 0;
      c.kLaunches[k] = 0;
      c.kRows[k] = 0;
    }
  });
}

int jbgpu_timing_kernels_stride(jbgpu_ctx* ctx, int32_t stride) {
  return guard(ctx, [&] {
    if (stride < 1) throw std::invalid_argument("iteration stride must be at least 1");
    ctx->c.kernelStride = stride;
  });
}

int jbgpu_timing_kernels_get(jbgpu_ctx* ctx, double* ms2, int64_t* launches2, int64_t* rows2) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    for (int k = 0; k < 4; k++) {
      if (ms2) ms2[k] = c.kMs[k];
      if (launches2) launches2[k] = c.kLaunches[k];
      if (rows2) rows2[k] = c.kRows[k];
    }
  });
}

int jbgpu_timing_get(jbgpu_ctx* ctx, double* ms, int64_t* launches, int64_t* steps) {
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    for (int r = 0; r < kTimerRegions; r++) {
      if (ms) ms[r] = c.ms[r];
      if (launches) launches[r] = c.launches[r];
    }
    if (steps) *steps = c.timedSteps;
  });
}

}  // extern "C"
