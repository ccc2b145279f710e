// internal.h — libjbgpu context, device buffers and the launchers shared by
// the kernel files.  Not part of the C ABI.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/jbgpu_format.h"

namespace jbg {

#define JB_CHECK(x)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// bumped by every device-buffer (re)allocation: a captured graph holds buffer addresses and is valid only while
// this stays what it was at capture
inline uint64_t gAllocEpoch = 0;

template <class T>
struct DBuf {  // device buffer owned by a context; freed with it
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void ensure(size_t k) {
    if (k <= n) return;
    if (p) {  // no kernel may still be reading the old buffer
      JB_CHECK(hipDeviceSynchronize());
      (void)hipFree(p);
    }
    size_t cap = k + k / 4 + 64;
    p = nullptr;
    JB_CHECK(hipMalloc((void**)&p, cap * sizeof(T)));
    n = cap;
    gAllocEpoch++;
  }
  void ensureKeep(size_t k) {  // as ensure, keeping the first n elements
    if (k <= n) return;
    size_t cap = k + k / 4 + 64;
    T* q = nullptr;
    JB_CHECK(hipMalloc((void**)&q, cap * sizeof(T)));
    if (p) {
      JB_CHECK(hipDeviceSynchronize());
      JB_CHECK(hipMemcpy(q, p, n * sizeof(T), hipMemcpyDeviceToDevice));
      (void)hipFree(p);
    }
    p = q;
    n = cap;
    gAllocEpoch++;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// pinned host buffer owned by a context: small device-to-host readbacks land here asynchronously (a copy into
// pageable memory holds the calling thread until it is done, one wait per copy)
template <typename T>
struct HBuf {
  T* p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  void ensure(size_t k) {  // callers synchronise before reusing the buffer
    if (k <= n) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    JB_CHECK(hipHostMalloc((void**)&p, k * sizeof(T), hipHostMallocDefault));
    n = k;
  }
};

// an instantiated graph, destroyed with its owner
struct GraphExec {
  hipGraphExec_t e = nullptr;
  std::vector<int64_t> key;  // what the capture baked in (see stepGraph)
  int stageAfter = 0, curAfter = 0;  // the host's Dbvt stage and pair buffer after one replay
  GraphExec() = default;
  GraphExec(const GraphExec&) = delete;
  GraphExec& operator=(const GraphExec&) = delete;
  ~GraphExec() { reset(); }
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
    key.clear();
  }
};

// device-side counts of a graph-captured step (int slots of Ctx::gc)
constexpr int kGcHuge = 0;       // huge proxies
constexpr int kGcLeaves = 1;     // tree leaves (proxies - huge), then the 32- and 1024-leaf chunk counts
constexpr int kGcChunk1 = 2;
constexpr int kGcChunk2 = 3;
constexpr int kGcPairs = 4;      // pairs found this step (0 once aborted)
constexpr int kGcAbort = 5;      // sticky: a step found more pairs than the capacity (or overflowed a stack)
constexpr int kGcDone = 6;       // steps completed since the host cleared it
constexpr int kGcActive = 7;     // active manifolds of the solve
constexpr int kGcPairCount = 8;  // pairs held by keys[0] and keys[1]
constexpr int kGcEpaOverflow = 10;  // sticky: an EPA job outgrew every workspace tier
constexpr int kGcWbErr = 11;        // sticky: error bits of the world-block solves
constexpr int kGcWords = 16;
constexpr int kResidStripes = 64;  // solver residual accumulators (per-wave stripes of (dd, nn) sums)

// EPA job workspace sizes (per queued pair)
constexpr int kEpaFaces = 1024;
constexpr int kEpaVerts = 300;
constexpr int kEpaHe = 160;
constexpr int kEpaFaceWords = 16;
constexpr int kEpaSlotWords = kEpaFaces * kEpaFaceWords + kEpaVerts * 6 + kEpaHe * 4 + 8;
// first-pass EPA workspace: small enough for several jobs per CU; jobs that outgrow it rerun with the large one
constexpr int kEpaSmallFaces = 192;
constexpr int kEpaSmallVerts = 96;
constexpr int kEpaSmallSlotWords = kEpaSmallFaces * kEpaFaceWords + kEpaSmallVerts * 6 + kEpaHe * 4 + 8;
constexpr int kEpaSmallGrid = 2048;  // first-pass EPA blocks (eight per CU)
constexpr int kEpaGrid = 512;  // EPA blocks (LDS workspace limits them to two per CU)
// last EPA tier (jobs that outgrow the large workspace): workspaces in global memory, one per block
constexpr int kEpaHugeFaces = 16384;
constexpr int kEpaHugeStack = 4096;
constexpr int kEpaHugeGrid = 16;
constexpr int kEpaHugeSlotWords = kEpaHugeFaces * kEpaFaceWords + kEpaVerts * 6 + kEpaHe * 4 + 8;
constexpr int kEpaHugeAuxInts = 64 + 2 * kEpaHugeStack + 5 * 6;
constexpr int kTimerRegions = 16;

// Constraint record: type (0 Point2PointConstraint, 1 HingeConstraint), body A, body B (-1: the fixed body), 0,
// then per type
//   point2point: pivot in A [4, 7), pivot in B (a world point for the fixed body) [7, 10), tau, damping, clamp
//   hinge: rbAFrame [4, 16) and rbBFrame [16, 28) (basis row-major, origin; B's third column flipped), lower and
//          upper limit, limit softness, bias factor, relaxation factor, angularOnly, enableAngularMotor, motor
//          target velocity, max motor impulse [28, 37)
constexpr int kJointWords = 40;
constexpr int kJointP2P = 0, kJointHinge = 1;
// per-solve state: 1 / diagonal of the three linear rows, appliedImpulse, the pivots' arms, pivotAInW -
// pivotBInW, both bodies' world inverse inertia [14, 32); hinge: linear axes [32, 41), world hinge axes of A and
// B [41, 47), kHinge, limit correction, limit sign, limit on, accumulated limit impulse [47, 52)
constexpr int kJointStateWords = 64;

// solver rows per manifold: 4 contacts x (normal row + 2 friction rows)
constexpr int kRowWords = 20;   // normal(3) rel1xn(3) rel2xn(3) angA(3) angB(3) jacInv pen rest fric applied
constexpr int kContactRowWords = 3 * kRowWords;
constexpr int kManifoldRowWords = 4 * kContactRowWords + 4;  // + [valid mask, bodyA, bodyB, _]

struct Stats {
  int64_t pairs = 0, manifoldsActive = 0, contacts = 0, rows = 0, colors = 0, epaCalls = 0, epaFails = 0;
  int64_t epaRetries = 0, epaHuge = 0;
  int64_t graphRedos = 0;
  int64_t iterations = 0;
  float residual = 0;
  int64_t huge = 0;
};

// LBVH over the bodies' current shape AABBs for world queries (rayTest); same layout as the broadphase tree
constexpr int kRecF4 = 5;  // float4 per traversal record: both children's boxes, then (child0, child1, last0, last1)
// float4 per 4-wide broadphase record: the boxes of a node's grandchildren (a leaf child stands for itself), then
// their ids and their last leaves as int4 (unused entries: last leaf -1)
constexpr int kRec4F4 = 10;
struct RayTree {
  DBuf<float> cand, aabb, leafBox, nodeBox, nodeRec, chunkBox;
  DBuf<int> isHuge, order, orderAlt, nodeChild, nodeFirst, nodeLast, hugeList, hugeListAlt, hugeWorld, hugeWorldAlt,
      hugeRange, leafExact;
  DBuf<uint64_t> morton, mortonAlt;
  DBuf<unsigned int> bounds;
  int n = 0, nn = 0, nHuge = 0;
  const int* sortedOrder = nullptr;
  // per query
  DBuf<int> count;
  DBuf<int64_t> offset;
  DBuf<uint64_t> keys, keysAlt;
  DBuf<float> rec, io;
  DBuf<int> ioInt;
  // convex sweeps: per query the origins of from and to, then the cast shape's AABB over the sweep (12 floats),
  // and the cast shape record
  DBuf<float> sweep, sweepShape;
  float sweepShapeHost[JBGPU_SHAPE_WORDS] = {};
};

struct Ctx {
  hipStream_t stream = nullptr;
  int device = 0;  // the HIP device current when the context was created
  std::string err;
  // configuration
  int bpMode = JBGPU_BP_DBVT;
  float wmin[3] = {-1000, -1000, -1000}, wmax[3] = {1000, 1000, 1000};
  float quant[3] = {0, 0, 0};
  uint32_t handleMask = 0xfffffffe, sentinel = 0x7fffffff, qmask = 0xffffffff;
  int numIterations = 10;
  int solverOrder = JBGPU_ORDER_COLOR;
  int deactivationDisabled = 1;
  float tolerance = 0.0f;
  int maxIterations = 100;
  int warmstart = 1;
  float erp = 0.2f, warmFactor = 0.85f, linearSlop = 0.0f;
  float gravity[3] = {0.0f, -9.81f, 0.0f};
  // world
  bool haveWorld = false;
  int nb = 0, ns = 0, nv = 0;
  DBuf<float> shapes, verts;
  DBuf<float> tr;     // nb*12
  DBuf<float> vel;    // nb*8  (lv xyz _, av xyz _)
  DBuf<float> svel;   // solver working velocities
  DBuf<float> props;  // nb*8  (invMass, invI xyz, friction, restitution, shape(int), flags(int))
  DBuf<float> grav;   // nb*4
  DBuf<int> act;      // activation
  DBuf<float> deact;  // deactivation time
  DBuf<int> world;    // world id per body/proxy
  // broadphase
  int np = 0;  // proxies
  int stageCurrent = 0;
  bool rotationPending = false;  // a setAabb pass ran the collide's stage rotation for the next pair query
  DBuf<int> stage;
  DBuf<float> leaf, aabb;  // np*6
  DBuf<int> filt;          // group | mask << 16
  DBuf<int> q;             // quantised edges np*6
  DBuf<float> cand;        // candidate boxes np*6 (float4 pairs: min xyz pad, max xyz pad -> np*8)
  DBuf<uint64_t> morton, mortonAlt;  // sort keys (world id above the Morton bits; 32-bit words when they fit)
  int worldBits = 0;                  // bits of the largest world id (0: a single world)
  int maxWorld = 0;                   // largest world id of the proxies
  std::vector<char> proxyDead;        // destroyed proxies (slots are not reused)
  DBuf<uint64_t> deltaKeys;           // pair deltas: compacted keys
  DBuf<int> deltaFlag;
  DBuf<int> order, orderAlt;  // sorted leaf -> proxy
  DBuf<int> isHuge;
  DBuf<int> hugeList, hugeListAlt, hugeWorld, hugeWorldAlt;
  int nHuge = 0;
  DBuf<float> leafBox;   // sorted leaves np*8
  DBuf<float> nodeBox;   // internal nodes np*8
  DBuf<float> nodeRec;   // traversal records np*20: both children's boxes, indices and last leaves
  DBuf<float> nodeRec4;  // 4-wide traversal records np*40 (kRec4F4 float4 per internal node)
  DBuf<int> hugeRange;   // per world: its range of the huge list
  RayTree rt;            // world queries
  // continuous collision detection: per body (motion threshold, swept sphere radius); the bodies with a
  // threshold other than 0 in index order; transforms before integration while it runs
  DBuf<float> ccd, trOld;
  DBuf<int> ccdList;
  std::vector<float> ccdHost;
  int nccd = 0;
  // graph-captured steps: the per-step counts on the device, their readback, one graph per (Dbvt stage, pair
  // buffer) parity
  DBuf<int> gc;
  HBuf<int> gcHost;
  GraphExec graphs[4];
  bool stepGraph = false;  // jbgpu_world_step replays captured graphs
  DBuf<int> leafExact;   // np*8: exact-test box, group/mask and proxy index of each sorted leaf
  DBuf<int> nodeChild;   // np*2
  DBuf<int> nodeLast;    // np
  DBuf<int> nodeFirst;   // np: internal node i covers sorted leaves [nodeFirst[i], nodeLast[i]]
  DBuf<float> chunkBox;  // unions of 32, 1024 and 32768 consecutive leaves (8 floats each)
  DBuf<int> flags;       // np
  DBuf<float> bounds;    // 8
  DBuf<uint64_t> rawKeys, rawKeysAlt;
  DBuf<uint64_t> pairStage;  // per traversal block: staged pair keys
  DBuf<unsigned long long> counter;  // misc device counters
  DBuf<int> primTmp;  // scratch of the device-wide scans, compactions and sorts (prim.hip)
  HBuf<int> primErrHost;  // the sorts' error word, read back after a step
  DBuf<int> primWide;
  int64_t pairCap = 0;
  // pairs and manifolds (double buffered for the remap)
  int cur = 0;
  int64_t npairs = 0, npairsOld = 0;
  DBuf<uint64_t> keys[2];
  DBuf<float> man[2];
  DBuf<int> ptype;
  DBuf<int> typeKey, typeKeyAlt, perm, permAlt;
  DBuf<unsigned long long> npStat;           // EPA queue length, failures, workspace overflows
  DBuf<int> binRange;                        // per pair type: first sorted position and count
  unsigned long long npStatHost[5] = {0, 0, 0, 0, 0};  // their copy, valid after the step's final synchronisation
  bool npStatPending = false;
  bool epaOverflow = false;  // the last narrowphase pass overflowed an EPA workspace
  std::vector<int> shapeTypeHost;  // host copies: type per shape and per body (canonical manifold headers)
  std::vector<int> bodyShapeType;
  std::vector<int> worldHost;       // host copies: world id and static flag per body (world-block tables)
  std::vector<char> bodyStaticHost;
  std::vector<float> bodyInvMassHost;
  // Point2Point constraints (DiscreteDynamicsWorld.addConstraint): records in creation order, their solve order
  // (greedy joint colours: joints of one colour share no body that a joint impulse moves) and per-solve state
  int nj = 0;
  std::vector<float> jointHost;       // kJointWords per joint
  DBuf<float> joints;                 // the same on the device
  DBuf<float> jointState;             // kJointStateWords per joint: 1 / diagonal, applied impulse, arms, inertia
  DBuf<int> jointOrder;               // joint indices, colour after colour (creation order inside a colour)
  DBuf<int> jointSolve;               // per joint: solved in the last solve (its island awake)
  std::vector<int> jointColourOff;    // colour k = jointOrder[off[k], off[k + 1])
  DBuf<int> jointColourOffDev;
  bool jointOrderReady = false;       // the order matches the joints and the bodies' static flags
  // batched worlds solved one workgroup per block of whole worlds (the whole solve in one launch)
  int wbBodies = 0;      // at most this many dynamic bodies per block (0: off)
  int wbThreads = 256;   // workgroup size of the block solve
  bool wbPipelined = false;  // rows of the next colour loaded while a colour is swept (more registers)
  bool wbReady = false;  // the tables below match wbBodies and the bodies
  int wbBlocks = 0, wbMaxBodies = 0;
  DBuf<int> wbBodyList;   // dynamic bodies, block after block
  DBuf<int> wbBodyStart;  // per block: first entry of wbBodyList (wbBlocks + 1 entries)
  DBuf<int> wbBodyBlock;  // per body: its block, -1 for static bodies
  DBuf<int> wbBodyLocal;  // per body: position in its block, -1 for static bodies
  DBuf<int> wbKey, wbKeyAlt, wbIdx, wbSlotStart, wbColour, wbSolveSlot;
  DBuf<int4> wbSlotInfo;  // per solve position: mask | first row << 4, bodies, packed LDS positions
  DBuf<int> wbInfo;       // [0] error flags, [1] most colours in a block
  HBuf<int> wbInfoHost;
  bool wbLastSolve = false;
  // ConstraintSolver seam: islands recorded between prepareSolve and allSolved
  bool solveBatchOpen = false;
  std::vector<int> islandManifolds;
  // the last solve's statistics: copied into pinned memory at its end, read after the step's final wait
  HBuf<float> solveRo;                // residual, iterations
  HBuf<double> solveRh;               // per-iteration (change, norm) sums
  HBuf<unsigned long long> solveRows; // rows in total, rows of the per-colour or block sweeps
  struct {
    bool pending = false, tolMode = false, blockSweeps = false, sampleK = false;
    int iters = 0, sampledIters = 0;  // sampledIters: iterations whose sweep launches were event-timed
    size_t kUsed = 0;
  } solveStat;
  DBuf<float> candC;  // contact candidates npairs*8
  DBuf<int> epaQueue;
  DBuf<int> epaRetry;  // EPA jobs that outgrew the first-pass workspace
  DBuf<int> epaRetryHuge;    // ... and then the large one
  int epaSmallFaces = kEpaSmallFaces, epaLargeFaces = kEpaFaces;  // face capacities the two LDS tiers use
  DBuf<float> epaHugeWork;   // the last tier's workspaces (kEpaHugeGrid x kEpaHugeSlotWords)
  DBuf<int> epaHugeAux;
  DBuf<float> epaState;  // npairs*12 GJK state for EPA merge
  // solver
  DBuf<float> rows;
  DBuf<int4> hdr;            // per colour slot: point mask | first row << 4, body A, body B, pair index
  DBuf<int> slotCount, slotStart;  // rows per slot and their exclusive prefix sum
  DBuf<float> nrow, frow;    // dense rows: normal row r, friction rows 2r and 2r+1
  DBuf<int2> activeBodies;   // per active manifold: body A, body B
  DBuf<int> colorStart;
  HBuf<int> batchHost;  // readback of the colour batches: count, first slots, boundary flags
  // learned warm start (16-64-64-64-3 MLP, bf16 weights, fp32 accumulate)
  bool mlpReady = false;
  bool useMlp = false;       // the model predicts the warm-start impulses of every solve
  bool mlpNewOnly = false;   // ... of the contacts without cached impulses only
  DBuf<float> mlpW;          // fp32 master parameters
  DBuf<uint16_t> mlpWb;      // bf16 MFMA operand fragments (forward W, backward W^T)
  DBuf<float> mlpNorm;       // input / output normalisation
  DBuf<float> mlpX, mlpY;    // inference / training samples
  DBuf<int> mlpIdx;          // training permutation
  DBuf<float> mlpG, mlpM, mlpV, mlpLoss;  // gradient, Adam moments, loss accumulator
  DBuf<uint16_t> mlpFrag;    // per training tile: activations and gradients, fragment-ready
  DBuf<int> mActive;  // active manifold list
  DBuf<int> island;      // island parent forest (root = smallest body index of the island)
  DBuf<int> islandFlag;  // per root: awake, inactive; then the convergence flag
  DBuf<unsigned long long> bodyPrio;
  DBuf<unsigned long long> bodyMask;
  DBuf<int> color, colorAlt;
  bool colourByDegree = false;  // Jones-Plassmann priorities largest degree first (false: hashed only)
  int recolourPasses = 0;      // iterated-greedy passes over the colour classes after the colouring
  DBuf<int> bodyDeg, mDeg;     // manifolds per dynamic body, per manifold the sum over its bodies
  DBuf<int> pend, pendAlt, pendFlag;  // uncoloured manifolds between colouring batches
  DBuf<uint64_t> colorKey, colorKeyAlt;
  DBuf<int> colorOrder, colorOrderAlt;
  std::vector<int64_t> colorOff;
  DBuf<int> colorOffDev;     // colour offsets on the device (tail sweeps, the per-body warm start)
  DBuf<int> warmSlotOf;      // per (colour batch, body): the body's slot in the batch, -1 for none
  // block-partitioned sweeps: dynamic bodies in blocks of blockSize consecutive Morton positions; manifolds
  // inside one block ("interior") are swept by one workgroup per block with the block's velocities in LDS,
  // the others ("boundary") by one launch per colour
  int blockSize = 0;                  // 0: every colour swept by its own launch
  const int* sortedOrder = nullptr;   // Morton-sorted proxy order of this step's broadphase
  int nBlocks = 0, interiorColors = 0;
  int64_t interiorEnd = 0;            // slots [0, interiorEnd) are interior
  DBuf<int> blockLocal;               // per body: position inside its block, -1 for static bodies
  DBuf<int> mPhase;                   // per active manifold: block index (interior) or -1 (boundary)
  DBuf<unsigned long long> slotKey, slotKeyAlt;
  DBuf<int> runStart, runEnd;         // per (interior colour, block): slot range
  DBuf<int2> slotLocal;               // per interior slot: LDS positions of its bodies (-1: static)
  DBuf<int> batchFlag;
  int tailThreshold = 1024;  // trailing colours of at most this many manifolds share one workgroup launch
  int headColors = 0;
  DBuf<double> resid;  // per iteration: dd, nn
  DBuf<int> done;
  HBuf<int> doneHost;   // tolerance mode: the done flag read back every few iterations
  int tolCheckEvery = 4;  // iterations between those reads (0: never; every iteration's kernels are launched)
  DBuf<float> learned;  // per-contact learned warm-start impulses (npairs*4*3)
  bool useLearned = false;
  int64_t learnedSlots = 0;
  DBuf<float> features;  // per-contact features npairs*4*16
  bool dumpFeatures = false;
  int64_t nActive = 0;
  int ncolors = 0;
  Stats st;
  // timing of kernel regions (HIP events on the context stream)
  bool timing = false;
  hipEvent_t evA[kTimerRegions] = {}, evB[kTimerRegions] = {};
  bool pending[kTimerRegions] = {};
  double ms[kTimerRegions] = {};
  int64_t launches[kTimerRegions] = {};
  int64_t timedSteps = 0;
  // sampled kernel timing of the solver sweeps (start/stop events per launch)
  int kernelEvery = 0;
  int kernelStride = 5;  // on a sampled solve, the sweeps of every kernelStride-th iteration are event-timed
  int64_t kernelSteps = 0;  // solves since the sampling period was set
  int colourRounds = 0;     // colouring rounds the last solve needed
  std::vector<hipEvent_t> kEv;  // event pool, two per timed launch
  std::vector<int> kEvKind;     // 0 normal rows, 1 friction rows
  double kMs[4] = {0, 0, 0, 0};  // 0/1: per-colour normal/friction launches, 2/3: block launches
  int64_t kLaunches[4] = {0, 0, 0, 0}, kRows[4] = {0, 0, 0, 0};
  DBuf<unsigned long long> rowCount;  // normal rows of the current solve
  std::vector<float> residHist;
};

// timer regions
enum {
  T_PREDICT = 0, T_SETAABB, T_TRAVERSE, T_GJK, T_MANIFOLD, T_SOLVE_SETUP, T_SOLVE_ITER, T_INTEGRATE, T_STEP,
  T_EPA, T_REMAP, T_MLP, T_BVH, T_KEYSORT, T_BIN, T_COLOR
};
inline void tBegin(Ctx& c, int r) {
  if (c.timing) (void)hipEventRecord(c.evA[r], c.stream);
}
inline void tEnd(Ctx& c, int r, int64_t nlaunch = 1) {
  if (!c.timing) return;
  (void)hipEventRecord(c.evB[r], c.stream);
  c.pending[r] = true;
  c.launches[r] += nlaunch;
}
inline void tCollect(Ctx& c) {  // after a stream synchronisation
  for (int r = 0; r < kTimerRegions; r++) {
    if (!c.pending[r]) continue;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, c.evA[r], c.evB[r]) == hipSuccess) c.ms[r] += t;
    c.pending[r] = false;
  }
}

// broadphase
void bpInitQuant(Ctx& c);
void bpUpdateFromBodies(Ctx& c);                        // A1 + A2 (+ collide stage rotation)
void bpUpdateFromAabbs(Ctx& c, const float* dAabb6);    // setAabb for all proxies from a device array
void bpFindPairs(Ctx& c);                                // A3-A7: sorted canonical keys in keys[cur]
void bpRemapManifolds(Ctx& c);                           // A5: pair persistence
void bpPairTypes(Ctx& c);                                // pair types of a pair set set by the caller
// narrowphase
void bpDestroyProxies(Ctx& c, int n, const int32_t* uids);  // BroadphaseInterface.destroyProxy
void rayBuildTree(Ctx& c);                               // LBVH over the current shape AABBs
void rayTest(Ctx& c, int n, const float* dFrom, const float* dTo, int16_t group, int16_t mask, const int* dWorld,
             int* dObj, float* dFrac, float* dNormal, float* dPoint);  // CollisionWorld.rayTest, closest hit
// CollisionWorld.convexSweepTest with a closest-hit callback for n queries of one convex shape (a shape record
// in host memory); from/to are 12-float transforms
void sweepTest(Ctx& c, int n, const float* castShape, const float* dFrom, const float* dTo, int16_t group,
               int16_t mask, const int* dWorld, int* dObj, float* dFrac, float* dNormal, float* dPoint);
void ccdClamp(Ctx& c, float dt, const int* abortFlag = nullptr, const int* dnp = nullptr);  // integrateTransforms' motion clamping
// graph-captured step (world.hip drives it): capacity-shaped launches with counts on the device
void graphFill(hipStream_t s, void* p, size_t words, uint32_t v);  // fill kernel (no memset node in a graph)
struct FillList {  // several fills in one launch
  static constexpr int kMax = 12;
  uint32_t* p[kMax];
  uint32_t n[kMax], v[kMax];
  int count = 0;
  void add(void* q, size_t words, uint32_t value) {
    if (count >= kMax) throw std::logic_error("fill list full");
    p[count] = (uint32_t*)q;
    n[count] = (uint32_t)words;
    v[count] = value;
    count++;
  }
};
void graphFillList(hipStream_t s, const FillList& f);
void bpGraphReserve(Ctx& c);
const uint64_t* bpStepGraph(Ctx& c);  // AABB update, tree, traversal, key sort; returns the sorted keys
void bpRemapGraph(Ctx& c, const uint64_t* sorted);
void npGraphReserve(Ctx& c);
void npDispatchGraph(Ctx& c);
void npStatsCopy(Ctx& c);  // the narrowphase statistics into pinned memory (after a graph launch)
void solverGraphReserve(Ctx& c);
void solverSolveGraph(Ctx& c, float dt);
void solverStatsCopy(Ctx& c);
void bpPairDeltas(Ctx& c, uint64_t* added, int64_t capAdded, int64_t* nAdded, uint64_t* removed, int64_t capRemoved,
                  int64_t* nRemoved);
void npDispatchAll(Ctx& c);                              // A8-A16
int64_t npCountContacts(Ctx& c);                         // contact points of the current manifolds
void npCollectStats(Ctx& c);                             // EPA statistics of the last narrowphase
void solverCollectStats(Ctx& c);                         // residual, rows and sweep timings of the last solve
// solver
// A18-A20; explicit: the manifolds to solve (pair indices, the islands a caller recorded) instead of the
// step's own selection (activity, islands, rows)
void solverSolve(Ctx& c, float dt, const std::vector<int>* explicitList = nullptr);
void solverGetImpulses(Ctx& c, float* out, int64_t nslots);  // per slot final (normal, lat1, lat2)
void jointsChanged(Ctx& c);                                  // joints or bodies changed: new solve order
void jointsImpulses(Ctx& c, float* out);                     // appliedImpulse per joint
int jointsOrderSolved(Ctx& c, int32_t* order, int cap);      // joints solved in the last solve, in order
// integration
void worldPredict(Ctx& c, float dt);
void worldIntegrate(Ctx& c, float dt);
void worldActivation(Ctx& c, float dt);

inline int gridFor(int64_t n, int block = 256) { return (int)((n + block - 1) / block); }

// device-wide primitives (prim.hip): the element count is read from *count (device memory) or, when count is
// null, taken from hostCount; launch shapes depend only on the capacity cap (count clamped to it), so the
// launches can be captured in a hipGraph.  scratch: primScratchInts(cap) ints of device memory.
int primTiles(int cap);
size_t primScratchInts(int cap, int keyBytes = 8);
// out[i] = in[0] + ... + in[i - 1] for i < n; *total (device, may be null) = sum of in[0, n)
void primExclusiveScan(hipStream_t s, const int* in, int* out, const int* count, int hostCount, int cap, int* scratch,
                       int* total);
// out[0, m) = the i < n with flag[i] != 0 in index order (vals[i] instead of i when vals is not null);
// *selectedCount (device) = m
void primSelectFlagged(hipStream_t s, const int* flag, const int* vals, int* out, const int* count, int hostCount,
                       int cap, int* scratch, int* selectedCount);
// stable LSD radix sort of the first n keys (and int values, when v0 is not null) by bits [bitLo, bitHi) rounded
// up to whole 8-bit digits (the keys' bits above that must be zero); returns true when the result is in (k1, v1)
template <class K>
bool primRadixSort(hipStream_t s, K* k0, int* v0, K* k1, int* v1, const int* count, int hostCount, int cap, int bitLo,
                   int bitHi, int* scratch);
// The step's forms (host-known count, scratch held by the context, work on the context stream).  A sort works on a
// pair of buffers and leaves its result in either: cur names it afterwards, alt the other one.
template <class T>
struct Db {
  T* cur;
  T* alt;
  Db(T* a, T* b) : cur(a), alt(b) {}
  T* Current() const { return cur; }
};
// stable sort of n keys (with int values) by bits [bitLo, bitHi); keys compare as unsigned words
template <class K>
void primSortPairs(Ctx& c, Db<K>& k, Db<int>& v, int n, int bitLo, int bitHi);
template <class K>
void primSortKeys(Ctx& c, Db<K>& k, int n, int bitLo, int bitHi);
// out[0, m) = the i < n with flag[i] != 0, in index order; *count (device) = m
void primSelectIndex(Ctx& c, const int* flag, int n, int* out, int* count);
// out[0, m) = vals[i] for the i < n with flag[i] != 0, in index order; *count (device) = m
template <class T>
void primSelectValues(Ctx& c, const T* vals, const int* flag, int n, T* out, int* count);
// out[i] = in[0] + ... + in[i - 1], i < n (the sums must fit in an int; primScanWide writes them as int64)
void primScan(Ctx& c, const int* in, int* out, int n);
void primScanWide(Ctx& c, const int* in, int64_t* out, int n);
// the same with the element count in device memory (*count, at most cap): launches shaped by cap only, as a
// captured graph needs; the sorted data ends in k.cur / v.cur (which buffer that is depends on cap and the bits)
void primReserve(Ctx& c, int cap);  // scratch for counts up to cap, allocated outside a capture
void primErrCopy(Ctx& c);   // the sorts' error word into pinned memory (asynchronous)
void primErrCheck(Ctx& c);  // after the stream's wait: throws if a sort's look-back gave up
template <class K>
void primSortPairsDev(Ctx& c, Db<K>& k, Db<int>& v, const int* count, int cap, int bitLo, int bitHi);
template <class K>
void primSortKeysDev(Ctx& c, Db<K>& k, const int* count, int cap, int bitLo, int bitHi);
void primSelectIndexDev(Ctx& c, const int* flag, const int* count, int cap, int* out, int* selected);
void primScanDev(Ctx& c, const int* in, int* out, const int* count, int cap);

}  // namespace jbg
