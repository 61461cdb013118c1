// bx_solve_quad.h -- velocity sweeps with eight lanes per island (component-parallel contact solve).
//
// The warm start and PGS sweeps of b3ContactSolver (contact_solver.cpp:272-486) are a serial chain
// per island: each contact reads the two body velocities the previous contact wrote. Inside one
// contact, though, almost all of the arithmetic is on 3-vectors: cross products, 3x3 inverse
// inertia times a vector, dot products, and v +=/-= m * P on body A and body B. Here a group of
// eight lanes solves one contact: lanes 0-3 of the group hold body A, lanes 4-7 body B, and lane
// 4*side + c holds component c (x, y, z; lane 3 of each quad is padding that computes along and
// never stores). Each lane evaluates exactly the binary32 operations the serial solver evaluates
// for its component, in the same order; the operands it needs from its siblings (the other
// components of a vector, body A's values on the B lanes) arrive by DPP lane permutes, which move
// bits and round nothing. Scalars (impulses, clamps, the friction cone) are evaluated on all eight
// lanes from the same inputs, so every lane holds the same impulse values, and each lane reloads
// in the next sweep only values it stored itself. The results are bit-identical to the serial
// solver's: the order of every rounded operation and of every sum is unchanged, and a subtraction
// a - b is evaluated as a + (-b), which IEEE defines identically (body A's updates use the negated
// inverse mass, and -1 * X for the inertia term, both exact negations).
//
// Body velocities live in an LDS slab per island (kSlab bodies); every joint row is solved by the
// group's first lane with the serial code of bx_solve_kernels.h. Islands that do not fit the slab
// (and the serial worlds' entries holding one) are left to k_solve_velocity, launched after this
// kernel for those islands only, so that this kernel's register count is that of its own path.
#pragma once
// (included by bx_step.hip inside namespace bx, after bx_solve_kernels.h)

// DPP moves. BX_DPP_MOV: every lane takes the value of the lane kCtrl names (a permutation inside
// each quad, where every source lane is valid; the compiler folds it into the instruction that
// uses it). BX_DPP_UPDATE: lanes of the banks in bankMask take it, the others keep `old`. The host
// build of the kernels defines both as its own group exchange.
#ifndef BX_DPP_MOV
#define BX_DPP_MOV(src, ctrl) __builtin_amdgcn_mov_dpp(src, ctrl, 0xF, 0xF, true)
#define BX_DPP_UPDATE(old, src, ctrl, bankMask) __builtin_amdgcn_update_dpp(old, src, ctrl, 0xF, bankMask, false)
#endif
// Lanes of a group exchange slab values through LDS in program order, which a wavefront's lockstep
// execution gives on the device; the host build, which runs the lanes one after another, defines
// BX_GROUP_SYNC as a wait for the group's eight lanes.
#ifndef BX_GROUP_SYNC
#define BX_GROUP_SYNC()
#endif
template <int kCtrl>
__device__ __forceinline__ scalar Dpp(scalar x) {
    return __builtin_bit_cast(scalar, BX_DPP_MOV(__builtin_bit_cast(int, x), kCtrl));
}
template <int kCtrl, int kBank>
__device__ __forceinline__ scalar DppBanks(scalar old, scalar x) {
    return __builtin_bit_cast(scalar, BX_DPP_UPDATE(__builtin_bit_cast(int, old), __builtin_bit_cast(int, x), kCtrl, kBank));
}
constexpr int kDppNext = 0xC9;        // quad_perm [1,2,0,3]: lane c reads component c+1 (mod 3)
constexpr int kDppNext2 = 0xD2;       // quad_perm [2,0,1,3]: component c+2 (mod 3)
constexpr int kDppBcastX = 0x00, kDppBcastY = 0x55, kDppBcastZ = 0xAA;
constexpr int kDppFromA = 0x114;      // row_shr:4 into the B banks: body A's value on the B lanes
constexpr int kDppToA = 0x104;        // row_shl:4 into the A banks: body B's value on the A lanes
constexpr int kBanksA = 0x5, kBanksB = 0xA;

__device__ __forceinline__ scalar QNext(scalar x) { return Dpp<kDppNext>(x); }
__device__ __forceinline__ scalar QNext2(scalar x) { return Dpp<kDppNext2>(x); }
// b3Dot over the quad's three components, ((x + y) + z), on every lane of the quad
__device__ __forceinline__ scalar QSum(scalar p) {
    const scalar s = Dpp<kDppBcastX>(p) + Dpp<kDppBcastY>(p);
    return s + Dpp<kDppBcastZ>(p);
}
// component c of Cross(a, b) = a[c+1] * b[c+2] - a[c+2] * b[c+1], with b's permuted components given
__device__ __forceinline__ scalar QCrossVar(scalar a, scalar bn, scalar bnn) { return QNext(a) * bnn - QNext2(a) * bn; }
// component c of Cross(r, p) for a lane-constant r (rn = r[c+1], rnn = r[c+2]) and a varying p
__device__ __forceinline__ scalar QCrossConst(scalar rn, scalar rnn, scalar p) { return rn * QNext2(p) - rnn * QNext(p); }
// component c of A * v = v.x * A.x + v.y * A.y + v.z * A.z (row c of A is i0, i1, i2)
__device__ __forceinline__ scalar QMatVec(scalar i0, scalar i1, scalar i2, scalar v) {
    const scalar s = Dpp<kDppBcastX>(v) * i0 + Dpp<kDppBcastY>(v) * i1;
    return s + Dpp<kDppBcastZ>(v) * i2;
}
__device__ __forceinline__ scalar FromA(scalar x) { return DppBanks<kDppFromA, kBanksB>(x, x); }
__device__ __forceinline__ scalar ToA(scalar x) { return DppBanks<kDppToA, kBanksA>(x, x); }

// A lane's role in its group.
struct QuadLane {
    u32 side;   // 0: body A, 1: body B
    u32 c;      // component 0..2 (3: padding lane)
    u32 cc;     // component this lane reads (the padding lane reads z)
    u32 cn, cnn;
    scalar sign;  // -1 on body A, +1 on body B
};

// Float offsets of the VelCon fields a lane reads.
constexpr u32 kVcInvMass = offsetof(VelCon, invMassA) / 4;
constexpr u32 kVcFriction = offsetof(VelCon, friction) / 4;
constexpr u32 kVcInvIA = offsetof(VelCon, invIA) / 4, kVcInvIB = offsetof(VelCon, invIB) / 4;
constexpr u32 kVcRA = offsetof(VelCon, rA) / 4, kVcRB = offsetof(VelCon, rB) / 4;
constexpr u32 kVcNormal = offsetof(VelCon, normal) / 4;
constexpr u32 kVcT1 = offsetof(VelCon, tangent1) / 4, kVcT2 = offsetof(VelCon, tangent2) / 4;
constexpr u32 kVcSpeeds = offsetof(VelCon, tangentSpeed1) / 4;  // tangentSpeed1/2, motorSpeed, motorMass
constexpr u32 kVcTangentMass = offsetof(VelCon, tangentMass) / 4;
constexpr u32 kVcImpulses = offsetof(VelCon, tangentImpulse) / 4;  // tangentImpulse x, y, motorImpulse
constexpr u32 kVcPoints = offsetof(VelCon, points) / 4;
constexpr u32 kVcPointStride = sizeof(VCPoint) / 4;
constexpr u32 kVcpRA = offsetof(VCPoint, rA) / 4, kVcpRB = offsetof(VCPoint, rB) / 4;
constexpr u32 kVcpNormal = offsetof(VCPoint, normal) / 4, kVcpNormalMass = offsetof(VCPoint, normalMass) / 4;
constexpr u32 kVcpNormalImpulse = offsetof(VCPoint, normalImpulse) / 4;
constexpr u32 kVcpBias = offsetof(VCPoint, velocityBias) / 4;
static_assert(offsetof(VelCon, tangentSpeed2) == offsetof(VelCon, tangentSpeed1) + 4 &&
                  offsetof(VelCon, motorSpeed) == offsetof(VelCon, tangentSpeed1) + 8 &&
                  offsetof(VelCon, motorMass) == offsetof(VelCon, tangentSpeed1) + 12 &&
                  offsetof(VelCon, motorImpulse) == offsetof(VelCon, tangentImpulse) + 8,
              "VelCon field order the quad loader assumes");

// One lane's share of a VelCon.
struct QCon {
    u32 ia, ib, pc;
    scalar m;              // inverse mass of the lane's body, negated on body A
    scalar i0, i1, i2;     // row cc of the body's inverse inertia
    scalar friction;
    scalar rn, rnn;        // the contact's arm (rA or rB) at components cn, cnn
    scalar n, t1, t2;      // normal and tangents at component cc
    scalar ts1, ts2, ms, mm;
    scalar kxx, kxy, kyx, kyy;  // tangentMass columns
    scalar tix, tiy, mi;   // tangent and motor impulses
    scalar prn[kMaxManifoldPoints], prnn[kMaxManifoldPoints], pn[kMaxManifoldPoints];
    scalar pm[kMaxManifoldPoints], pi[kMaxManifoldPoints], pb[kMaxManifoldPoints];
};

__device__ __forceinline__ void QLoad(QCon& q, const VelCon* src, const QuadLane& L) {
    const scalar* f = reinterpret_cast<const scalar*>(src);
    const u32* u = reinterpret_cast<const u32*>(src);
    q.ia = u[0];
    q.ib = u[1];
    q.pc = u[2];
    const scalar m = f[kVcInvMass + L.side];
    q.m = L.side ? m : -m;
    const u32 I = L.side ? kVcInvIB : kVcInvIA;
    q.i0 = f[I + L.cc];
    q.i1 = f[I + 3 + L.cc];
    q.i2 = f[I + 6 + L.cc];
    q.friction = f[kVcFriction];
    const u32 r = L.side ? kVcRB : kVcRA;
    q.rn = f[r + L.cn];
    q.rnn = f[r + L.cnn];
    q.n = f[kVcNormal + L.cc];
    q.t1 = f[kVcT1 + L.cc];
    q.t2 = f[kVcT2 + L.cc];
    q.ts1 = f[kVcSpeeds];
    q.ts2 = f[kVcSpeeds + 1];
    q.ms = f[kVcSpeeds + 2];
    q.mm = f[kVcSpeeds + 3];
    q.kxx = f[kVcTangentMass];
    q.kxy = f[kVcTangentMass + 1];
    q.kyx = f[kVcTangentMass + 2];
    q.kyy = f[kVcTangentMass + 3];
    q.tix = f[kVcImpulses];
    q.tiy = f[kVcImpulses + 1];
    q.mi = f[kVcImpulses + 2];
    const u32 pr = L.side ? kVcpRB : kVcpRA;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        const scalar* p = f + kVcPoints + k * kVcPointStride;
        q.prn[k] = p[pr + L.cn];
        q.prnn[k] = p[pr + L.cnn];
        q.pn[k] = p[kVcpNormal + L.cc];
        q.pm[k] = p[kVcpNormalMass];
        q.pi[k] = p[kVcpNormalImpulse];
        q.pb[k] = p[kVcpBias];
    }
}

// Body A's value of x on the B lanes, combined as b - a_from_A.
__device__ __forceinline__ scalar SubA(scalar b, scalar x) { return b - FromA(x); }

// v, w: the lane's component of its body's velocities. contact_solver.cpp:272-326.
__device__ __forceinline__ void QWarmStart(const QCon& q, scalar& v, scalar& w, const QuadLane& L) {
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= q.pc) break;
        const scalar P = q.pi[k] * q.pn[k];
        v = v + q.m * P;
        w = w + L.sign * QMatVec(q.i0, q.i1, q.i2, QCrossConst(q.prn[k], q.prnn[k], P));
    }
    if (q.pc > 0) {
        const scalar P12 = q.tix * q.t1 + q.tiy * q.t2;
        const scalar P3 = q.mi * q.n;
        v = v + q.m * P12;
        w = w + L.sign * QMatVec(q.i0, q.i1, q.i2, QCrossConst(q.rn, q.rnn, P12) + P3);
    }
}

// contact_solver.cpp:328-486; the impulses in q are updated on every lane.
__device__ __forceinline__ void QSolve(QCon& q, scalar& v, scalar& w, const QuadLane& L) {
    scalar normalImpulse = 0;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= q.pc) break;
        // dv = vB + Cross(wB, rB) - vA - Cross(wA, rA), on the B lanes
        const scalar cr = QCrossVar(w, q.prn[k], q.prnn[k]);
        const scalar dv = SubA(SubA(v + cr, v), cr);
        const scalar Cdot = ToA(QSum(q.pn[k] * dv));
        scalar impulse = -q.pm[k] * (Cdot - q.pb[k]);
        const scalar oldImpulse = q.pi[k];
        q.pi[k] = Max(q.pi[k] + impulse, scalar(0));
        impulse = q.pi[k] - oldImpulse;
        const scalar P = impulse * q.pn[k];
        v = v + q.m * P;
        w = w + L.sign * QMatVec(q.i0, q.i1, q.i2, QCrossConst(q.prn[k], q.prnn[k], P));
        normalImpulse += q.pi[k];
    }
    if (q.pc > 0) {
        {
            const scalar cr = QCrossVar(w, q.rn, q.rnn);
            const scalar dv = SubA(SubA(v + cr, v), cr);
            const scalar cx = ToA(QSum(dv * q.t1) - q.ts1);
            const scalar cy = ToA(QSum(dv * q.t2) - q.ts2);
            // tangentMass * -Cdot = (-cx) * K.x + (-cy) * K.y
            scalar ix = -cx * q.kxx + -cy * q.kyx;
            scalar iy = -cx * q.kxy + -cy * q.kyy;
            const scalar ox = q.tix, oy = q.tiy;
            q.tix += ix;
            q.tiy += iy;
            const scalar maxImpulse = q.friction * normalImpulse;
            if (q.tix * q.tix + q.tiy * q.tiy > maxImpulse * maxImpulse) {
                const scalar len = Sqrt(q.tix * q.tix + q.tiy * q.tiy);  // b3Vec2::Normalize
                if (len > kEpsilon) {
                    q.tix /= len;
                    q.tiy /= len;
                }
                q.tix *= maxImpulse;
                q.tiy *= maxImpulse;
            }
            ix = q.tix - ox;
            iy = q.tiy - oy;
            const scalar P = ix * q.t1 + iy * q.t2;
            v = v + q.m * P;
            w = w + L.sign * QMatVec(q.i0, q.i1, q.i2, QCrossConst(q.rn, q.rnn, P));
        }
        {
            const scalar Cdot = ToA(QSum(q.n * SubA(w, w)) - q.ms);
            scalar impulse = -q.mm * Cdot;
            const scalar oldImpulse = q.mi;
            const scalar maxImpulse = q.friction * normalImpulse;
            q.mi = Clamp(q.mi + impulse, -maxImpulse, maxImpulse);
            impulse = q.mi - oldImpulse;
            const scalar P = impulse * q.n;
            w = w + L.sign * QMatVec(q.i0, q.i1, q.i2, P);
        }
    }
}

// Writes the impulses a sweep accumulated (every lane stores the same values).
__device__ __forceinline__ void QStoreImpulses(const QCon& q, VelCon* dst) {
    scalar* f = reinterpret_cast<scalar*>(dst);
    f[kVcImpulses] = q.tix;
    f[kVcImpulses + 1] = q.tiy;
    f[kVcImpulses + 2] = q.mi;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k)
        if (k < q.pc) f[kVcPoints + k * kVcPointStride + kVcpNormalImpulse] = q.pi[k];
}

// One contact of a sweep on the group's slab (6 floats per body: v, w); the constraint at `next` is
// requested into `ahead`.
template <bool kSolve>
__device__ __forceinline__ void QStep(QCon& cur, QCon& ahead, const VelCon* next, VelCon* self, scalar* slab,
                                      const QuadLane& L) {
    BX_GROUP_SYNC();  // the previous contact's velocities are in the slab
    const u32 b = (L.side ? cur.ib : cur.ia) * 6u + L.cc;
    scalar v = slab[b], w = slab[b + 3];
    QLoad(ahead, next, L);
    if (kSolve) QSolve(cur, v, w, L);
    else QWarmStart(cur, v, w, L);
    if (L.c < 3) {
        slab[b] = v;
        slab[b + 3] = w;
    }
    if (kSolve) QStoreImpulses(cur, self);
}

// Constraints a sweep requests ahead of the one it solves: each lane holds kQuadAhead + 1 register
// buffers, so a constraint's loads have kQuadAhead solves to arrive in.
#ifndef BX_QUAD_AHEAD
#define BX_QUAD_AHEAD 2
#endif
constexpr u32 kQuadAhead = BX_QUAD_AHEAD;

template <bool kSolve>
__device__ __forceinline__ void QSweep(VelCon* vcs, u32 nc, scalar* slab, const QuadLane& L) {
    // the buffers rotate: the constraint kQuadAhead places ahead is requested into the buffer of the
    // one solved last, while the current one is solved
    constexpr u32 kBuf = kQuadAhead + 1;
    QCon q[kBuf];
#pragma unroll
    for (u32 s = 0; s < kQuadAhead; ++s) QLoad(q[s], vcs + Min(s, nc - 1), L);
    for (u32 k = 0;;) {
#pragma unroll
        for (u32 s = 0; s < kBuf; ++s) {
            QStep<kSolve>(q[s], q[(s + kQuadAhead) % kBuf], vcs + Min(k + kQuadAhead, nc - 1), vcs + k, slab, L);
            if (++k == nc) return;
        }
    }
}

// kGroups islands per 64-lane workgroup, eight lanes each; velocities of islands of at most kSlab
// bodies in LDS.
template <u32 kGroups, u32 kSlab>
__global__ void __launch_bounds__(64) k_solve_velocity_quad(BatchView B, StepParams sp) {
    static_assert(kGroups * 8 <= 64, "eight lanes per island");
    __shared__ scalar slabs[kGroups][kSlab * 6];
    const u32 lane = threadIdx.x;
    const u32 g = lane >> 3;
    if (g >= kGroups) return;
    QuadLane L;
    L.side = (lane >> 2) & 1u;
    L.c = lane & 3u;
    L.cc = L.c < 3 ? L.c : 2u;
    L.cn = L.cc == 2 ? 0u : L.cc + 1;
    L.cnn = L.cn == 2 ? 0u : L.cn + 1;
    L.sign = L.side ? scalar(1) : scalar(-1);
    const u32 inGroup = lane & 7u;
    scalar* slab = slabs[g];
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * kGroups + g; i < count; i += gridDim.x * kGroups) {
        const uint2 e = B.islandList[i];
        BX_WORLD_HOOK(B, e.x);
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const bool warm = (wi.flags & 0x2u) == 0;
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        if (EntryLargestIsland(B, wi, first, last) > kSlab) continue;  // k_solve_velocity's
        for (u32 k = first; k < last; ++k) {
            const IslandInfo isl = B.islands[wi.bodyBase + k];
            VelBody* gvel = B.velBodies + wi.islandBodyBase + isl.bodyStart;
            VelCon* vcs = B.velCons + wi.conBase + isl.contactStart;
            JointCon* jcs = B.jointCons + wi.jointBase + isl.jointStart;
            const u32 nc = isl.contactCount, nj = isl.jointCount;
            const scalar* gf = reinterpret_cast<const scalar*>(gvel);
            BX_GROUP_SYNC();  // the previous island's velocities have left the slab
            for (u32 t = inGroup; t < isl.bodyCount * 6u; t += 8) slab[t] = gf[t];
            BX_GROUP_SYNC();
            VelBody* vb = reinterpret_cast<VelBody*>(slab);
            // island.cpp:154-181: contact warm start, joint warm start, then per iteration all
            // joints followed by all contacts
            // the group's first lane solves the joints on velocities the other lanes' contact
            // sweeps wrote; each sweep waits for it before it reads the slab
            if (warm && nc > 0) QSweep<false>(vcs, nc, slab, L);
            BX_GROUP_SYNC();
            if (warm && nj > 0 && inGroup == 0)
                for (u32 j = 0; j < nj; ++j) WarmStartJoint(jcs[j], vb);
            for (u32 it = 0; it < sp.velocityIterations && (nc > 0 || nj > 0); ++it) {
                BX_GROUP_SYNC();
                if (nj > 0 && inGroup == 0)
                    for (u32 j = 0; j < nj; ++j) SolveJointVelocity(jcs[j], vb, sp.dt);
                if (nc > 0) QSweep<true>(vcs, nc, slab, L);
            }
            BX_GROUP_SYNC();
            scalar* gout = reinterpret_cast<scalar*>(gvel);
            for (u32 t = inGroup; t < isl.bodyCount * 6u; t += 8) gout[t] = slab[t];
        }
    }
}


// ---- Position sweeps with eight lanes per island (contact_solver.cpp:517-624) ----
//
// The NGS position solve of one contact point rotates the local center, normal and points by the
// bodies' quaternions, solves one scalar impulse, moves the centers, updates the quaternions by
// their derivative, renormalizes them and rotates the local inverse inertia into the world frame
// again. The same eight lanes hold each body's vectors by component; a quaternion's vector part sits
// on lanes 0-2 of the body's quad and its scalar part on lane 3. Each lane evaluates the serial
// solver's operations for its component in the same order, the inverse inertia's row c on lane c.
// The island's poses (x, q) live in an LDS slab; joints are solved by the group's first lane with the
// serial code.

constexpr int kDppBcastW = 0xFF;  // quad_perm [3,3,3,3]: the quaternion's scalar part

// PosCon float offsets.
constexpr u32 kPcInvMass = offsetof(PosCon, invMassA) / 4, kPcRadius = offsetof(PosCon, radiusA) / 4;
constexpr u32 kPcCenterA = offsetof(PosCon, localCenterA) / 4, kPcCenterB = offsetof(PosCon, localCenterB) / 4;
constexpr u32 kPcInvIA = offsetof(PosCon, localInvIA) / 4, kPcInvIB = offsetof(PosCon, localInvIB) / 4;
constexpr u32 kPcNormalA = offsetof(PosCon, localNormalA) / 4;
constexpr u32 kPcPointA = offsetof(PosCon, localPointA) / 4, kPcPointB = offsetof(PosCon, localPointB) / 4;
static_assert(offsetof(PosCon, invMassB) == offsetof(PosCon, invMassA) + 4 &&
                  offsetof(PosCon, radiusB) == offsetof(PosCon, radiusA) + 4,
              "PosCon field order the quad loader assumes");
// PosBody float offsets: x 0-2, q.v 3-5, q.s 6, inverse inertia columns 7-9, 10-12, 13-15.
static_assert(sizeof(PosBody) == 64 && offsetof(PosBody, q) == 12 && offsetof(PosBody, I) == 28,
              "PosBody layout the position slab assumes");

// One lane's share of a PosCon.
struct PCon {
    u32 ia, ib, pc;
    scalar mA, mB, radA, radB;
    scalar m;                 // the lane's body's inverse mass, negated on body A
    scalar lc, lcn, lcnn;     // the body's local center at components cc, cn, cnn
    scalar L[9];              // the body's local inverse inertia (columns x, y, z)
    // per point: body A's lanes rotate the local normal, body B's its local point (components cc,
    // cn, cnn), then both quads rotate body A's local point
    scalar v2[kMaxManifoldPoints], v2n[kMaxManifoldPoints], v2nn[kMaxManifoldPoints];
    scalar v3[kMaxManifoldPoints], v3n[kMaxManifoldPoints], v3nn[kMaxManifoldPoints];
};

__device__ __forceinline__ void PLoad(PCon& p, const PosCon* src, const QuadLane& L) {
    const scalar* f = reinterpret_cast<const scalar*>(src);
    const u32* u = reinterpret_cast<const u32*>(src);
    p.ia = u[0];
    p.ib = u[1];
    p.pc = u[2];
    p.mA = f[kPcInvMass];
    p.mB = f[kPcInvMass + 1];
    p.radA = f[kPcRadius];
    p.radB = f[kPcRadius + 1];
    p.m = L.side ? p.mB : -p.mA;
    const u32 lc = L.side ? kPcCenterB : kPcCenterA;
    p.lc = f[lc + L.cc];
    p.lcn = f[lc + L.cn];
    p.lcnn = f[lc + L.cnn];
    const u32 I = L.side ? kPcInvIB : kPcInvIA;
#pragma unroll
    for (u32 j = 0; j < 9; ++j) p.L[j] = f[I + j];
    const u32 v2 = L.side ? kPcPointB : kPcNormalA;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        p.v2[k] = f[v2 + 3 * k + L.cc];
        p.v2n[k] = f[v2 + 3 * k + L.cn];
        p.v2nn[k] = f[v2 + 3 * k + L.cnn];
        p.v3[k] = f[kPcPointA + 3 * k + L.cc];
        p.v3n[k] = f[kPcPointA + 3 * k + L.cn];
        p.v3nn[k] = f[kPcPointA + 3 * k + L.cnn];
    }
}

// component c of Mul(q, v) = v + q.s * t + Cross(q.v, t), t = 2 * Cross(q.v, v) (quat.h:270-274), for
// a lane-constant v (components cc, cn, cnn); q holds the quaternion (q.v on lanes 0-2, q.s on 3)
__device__ __forceinline__ scalar QRotate(scalar q, scalar qs, scalar v, scalar vn, scalar vnn) {
    const scalar t = scalar(2) * QCrossVar(q, vn, vnn);
    return (v + qs * t) + QCrossVar(q, QNext(t), QNext2(t));
}

// Lane c's row of RotateToFrame(L, q) = R * L * Transpose(R), R = q.GetRotationMatrix() (quat.h,
// mat33.h): i0, i1, i2 = columns x, y, z of the result at component c.
__device__ __forceinline__ void QRotateToFrame(const scalar* L, scalar q, const QuadLane& Ln, scalar& i0, scalar& i1,
                                               scalar& i2) {
    const scalar x = Dpp<kDppBcastX>(q), y = Dpp<kDppBcastY>(q), z = Dpp<kDppBcastZ>(q), w = Dpp<kDppBcastW>(q);
    const scalar x2 = x + x, y2 = y + y, z2 = z + z;
    const scalar xx = x * x2, xy = x * y2, xz = x * z2;
    const scalar yy = y * y2, yz = y * z2, zz = z * z2;
    const scalar wx = w * x2, wy = w * y2, wz = w * z2;
    // R's columns
    const scalar Rx0 = scalar(1) - (yy + zz), Rx1 = xy + wz, Rx2 = xz - wy;
    const scalar Ry0 = xy - wz, Ry1 = scalar(1) - (xx + zz), Ry2 = yz + wx;
    const scalar Rz0 = xz + wy, Rz1 = yz - wx, Rz2 = scalar(1) - (xx + yy);
    const u32 c = Ln.cc;
    const scalar rx = c == 0 ? Rx0 : (c == 1 ? Rx1 : Rx2);
    const scalar ry = c == 0 ? Ry0 : (c == 1 ? Ry1 : Ry2);
    const scalar rz = c == 0 ? Rz0 : (c == 1 ? Rz1 : Rz2);
    // (R * L) column j at component c: (L_j.x * R.x + L_j.y * R.y) + L_j.z * R.z
    const scalar a0 = (L[0] * rx + L[1] * ry) + L[2] * rz;
    const scalar a1 = (L[3] * rx + L[4] * ry) + L[5] * rz;
    const scalar a2 = (L[6] * rx + L[7] * ry) + L[8] * rz;
    // (R L) * Transpose(R) column k at component c: (R.x[k] * RL.x + R.y[k] * RL.y) + R.z[k] * RL.z
    i0 = (Rx0 * a0 + Ry0 * a1) + Rz0 * a2;
    i1 = (Rx1 * a0 + Ry1 * a1) + Rz1 * a2;
    i2 = (Rx2 * a0 + Ry2 * a1) + Rz2 * a2;
}

// One contact of a position sweep (contact_solver.cpp:524-621); x, q, i0-i2 are the lane's share of
// its body's pose; returns the running minimum separation (the same on every lane).
__device__ __forceinline__ scalar QSolvePosition(const PCon& p, scalar& x, scalar& q, scalar& i0, scalar& i1,
                                                 scalar& i2, scalar minSeparation, const QuadLane& L) {
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= p.pc) break;
        const scalar qs = Dpp<kDppBcastW>(q);
        const scalar xt = x - QRotate(q, qs, p.lc, p.lcn, p.lcnn);  // xfA / xfB translations
        const scalar r2 = QRotate(q, qs, p.v2[k], p.v2n[k], p.v2nn[k]);  // A: normal; B: Mul(qB, pointB)
        const scalar c1 = QRotate(q, qs, p.v3[k], p.v3n[k], p.v3nn[k]) + xt;  // A: Mul(xfA, pointA)
        const scalar c2 = r2 + xt;                                  // B: Mul(xfB, pointB)
        const scalar normal = FromA(r2);
        // separation = Dot(c2 - c1, normal) - radiusA - radiusB, on the B lanes
        const scalar separation = ToA((QSum((c2 - FromA(c1)) * normal) - p.radA) - p.radB);
        minSeparation = Min(minSeparation, separation);
        const scalar C = Clamp(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, scalar(0));
        const scalar r = ToA(c2) - x;                               // rA = point - cA, rB = point - cB
        const scalar rn = QCrossVar(r, QNext(normal), QNext2(normal));
        const scalar d = QSum(rn * QMatVec(i0, i1, i2, rn));       // Dot(rn, I * rn) of the lane's body
        const scalar K = ((p.mA + p.mB) + FromA(d)) + ToA(d);
        const scalar impulse = K > scalar(0) ? -C / K : scalar(0);
        const scalar P = impulse * normal;
        x = x + p.m * P;                                            // cA -= mA * P, cB += mB * P
        // q -/+= Derivative(q, I * Cross(r, P)) = (0.5 * Quat(omega, 0)) * q, the product
        // Quat(Cross(a.v, q.v) + a.s * q.v + q.s * a.v, a.s * q.s - Dot(a.v, q.v)) with a = (0.5 omega, 0.5 * 0)
        const scalar av = scalar(0.5) * QMatVec(i0, i1, i2, QCrossVar(r, QNext(P), QNext2(P)));
        const scalar as = scalar(0.5) * scalar(0);
        const scalar dv = (QCrossVar(av, QNext(q), QNext2(q)) + as * q) + qs * av;
        const scalar ds = as * q - QSum(av * q);
        q = q + L.sign * (L.c < 3 ? dv : ds);
        // q.Normalize(): the vector part multiplies by the reciprocal, the scalar part divides
        const scalar qq = q * q;
        const scalar len = Sqrt(QSum(qq) + Dpp<kDppBcastW>(qq));
        if (len > kEpsilon) {
            const scalar inv = scalar(1) / len;
            q = L.c < 3 ? q * inv : q / len;
        }
        QRotateToFrame(p.L, q, L, i0, i1, i2);
    }
    return minSeparation;
}

// The pose slab holds 8 words per body: x (0-2), q (3-6) and a flag (7). The world inverse inertia is
// not stored: after any position update of a body it is RotateToFrame(invI, q) with the body's own
// local inverse inertia (contact_solver.cpp:580-621, revolute_joint.cpp), and before the first one it
// is that too for every body the position integration moved (island.cpp:190-229 sets it so); only a
// static body starts from its stored matrix, which the flag marks until the body is first updated.
constexpr u32 kPoseWords = 8;
constexpr u32 kPoseFromGlobal = 1;  // flag: the inverse inertia is the island's stored one

// The lane's row of the body's world inverse inertia at the start of a contact.
__device__ __forceinline__ void PoseInertia(const scalar* L, scalar q, u32 flag, const PosBody* gpos, u32 body,
                                            const QuadLane& Ln, scalar& i0, scalar& i1, scalar& i2) {
    QRotateToFrame(L, q, Ln, i0, i1, i2);
    if (flag & kPoseFromGlobal) {
        const Mat33& I = gpos[body].I;
        i0 = I.x[Ln.cc];
        i1 = I.y[Ln.cc];
        i2 = I.z[Ln.cc];
    }
}

// One contact of a position sweep on the group's slab; the constraint at `next` is requested into
// `ahead`.
__device__ __forceinline__ scalar PStep(const PCon& cur, PCon& ahead, const PosCon* next, scalar* slab,
                                        const PosBody* gpos, scalar minSeparation, const QuadLane& L) {
    BX_GROUP_SYNC();  // the previous contact's poses are in the slab
    const u32 body = L.side ? cur.ib : cur.ia;
    const u32 b = body * kPoseWords;
    scalar x = slab[b + L.cc], q = slab[b + 3 + L.c];
    const u32 flag = __builtin_bit_cast(u32, slab[b + 7]);
    PLoad(ahead, next, L);
    scalar i0, i1, i2;
    PoseInertia(cur.L, q, flag, gpos, body, L, i0, i1, i2);
    minSeparation = QSolvePosition(cur, x, q, i0, i1, i2, minSeparation, L);
    if (L.c < 3) slab[b + L.c] = x;
    slab[b + 3 + L.c] = q;
    if (L.c == 0 && cur.pc > 0) slab[b + 7] = __builtin_bit_cast(scalar, 0u);  // updated: I follows q now
    return minSeparation;
}

// A pose of the slab as the serial joint code reads it.
struct SlabPose {
    Vec3 x;
    Quat q;
    Mat33 I;
};

__device__ __forceinline__ SlabPose LoadPose(const scalar* slab, const PosBody* gpos, u32 body, const Mat33& invI) {
    const scalar* r = slab + body * kPoseWords;
    SlabPose p;
    p.x = Vec3(r[0], r[1], r[2]);
    p.q = Quat(Vec3(r[3], r[4], r[5]), r[6]);
    p.I = (__builtin_bit_cast(u32, r[7]) & kPoseFromGlobal) ? gpos[body].I : RotateToFrame(invI, p.q);
    return p;
}

__device__ __forceinline__ void StorePose(scalar* slab, u32 body, const SlabPose& p) {
    scalar* r = slab + body * kPoseWords;
    r[0] = p.x.x;
    r[1] = p.x.y;
    r[2] = p.x.z;
    r[3] = p.q.v.x;
    r[4] = p.q.v.y;
    r[5] = p.q.v.z;
    r[6] = p.q.s;
    r[7] = __builtin_bit_cast(scalar, 0u);
}

// kGroups islands per 64-lane workgroup, eight lanes each; poses of islands of at most kSlab bodies in
// LDS (larger islands, and serial-world entries holding one, are k_solve_position's).
template <u32 kGroups, u32 kSlab>
__global__ void __launch_bounds__(64) k_solve_position_quad(BatchView B, StepParams sp) {
    static_assert(kGroups * 8 <= 64, "eight lanes per island");
    __shared__ scalar slabs[kGroups][kSlab * kPoseWords];
    __shared__ u32 jointFlags[kGroups];
    const u32 lane = threadIdx.x;
    const u32 g = lane >> 3;
    if (g >= kGroups) return;
    QuadLane L;
    L.side = (lane >> 2) & 1u;
    L.c = lane & 3u;
    L.cc = L.c < 3 ? L.c : 2u;
    L.cn = L.cc == 2 ? 0u : L.cc + 1;
    L.cnn = L.cn == 2 ? 0u : L.cn + 1;
    L.sign = L.side ? scalar(1) : scalar(-1);
    const u32 inGroup = lane & 7u;
    scalar* slab = slabs[g];
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * kGroups + g; i < count; i += gridDim.x * kGroups) {
        const uint2 e = B.islandList[i];
        BX_WORLD_HOOK(B, e.x);
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        if (EntryLargestIsland(B, wi, first, last) > kSlab) continue;  // k_solve_position's
        const Body* bodies = B.bodies + wi.bodyBase;
        for (u32 k = first; k < last; ++k) {
            IslandInfo& isl = B.islands[wi.bodyBase + k];
            PosBody* gpos = B.posBodies + wi.islandBodyBase + isl.bodyStart;
            const u32* ib = B.islandBodies + wi.islandBodyBase + isl.bodyStart;
            const PosCon* pcs = B.posCons + wi.conBase + isl.contactStart;
            const JointCon* jcs = B.jointCons + wi.jointBase + isl.jointStart;
            const u32 nc = isl.contactCount, nj = isl.jointCount;
            BX_GROUP_SYNC();  // the previous island's poses have left the slab
            for (u32 t = inGroup; t < isl.bodyCount; t += 8) {
                const PosBody& p = gpos[t];
                scalar* r = slab + t * kPoseWords;
                r[0] = p.x.x;
                r[1] = p.x.y;
                r[2] = p.x.z;
                r[3] = p.q.v.x;
                r[4] = p.q.v.y;
                r[5] = p.q.v.z;
                r[6] = p.q.s;
                r[7] = __builtin_bit_cast(scalar, bodies[ib[t]].type == kStatic ? kPoseFromGlobal : 0u);
            }
            BX_GROUP_SYNC();
            // island.cpp:231-248 / PositionSweeps: contacts, then joints, until both are solved
            bool solved = false;
            for (u32 it = 0; it < sp.positionIterations; ++it) {
                scalar minSeparation = 0;
                if (nc > 0) {
                    PCon a, b;
                    PLoad(a, pcs, L);
                    for (u32 c = 0;;) {
                        minSeparation = PStep(a, b, pcs + Min(c + 1, nc - 1), slab, gpos, minSeparation, L);
                        if (++c == nc) break;
                        minSeparation = PStep(b, a, pcs + Min(c + 1, nc - 1), slab, gpos, minSeparation, L);
                        if (++c == nc) break;
                    }
                }
                const bool contactsSolved = minSeparation >= scalar(-3) * kLinearSlop;
                BX_GROUP_SYNC();  // the contacts' poses are in the slab
                if (inGroup == 0) {
                    bool jointsSolved = true;  // joint_solver.cpp:55-64
                    for (u32 j = 0; j < nj; ++j) {
                        JointCon jc = jcs[j];
                        SlabPose two[2] = {LoadPose(slab, gpos, jc.indexA, jc.localInvIA),
                                           LoadPose(slab, gpos, jc.indexB, jc.localInvIB)};
                        const u32 a = jc.indexA, bb = jc.indexB;
                        jc.indexA = 0;
                        jc.indexB = 1;
                        const bool ok = SolveJointPosition(jc, two);
                        StorePose(slab, a, two[0]);
                        StorePose(slab, bb, two[1]);
                        jointsSolved = jointsSolved && ok;
                    }
                    jointFlags[g] = jointsSolved ? 1u : 0u;
                }
                BX_GROUP_SYNC();  // the joints' poses and flag are in LDS
                if (contactsSolved && jointFlags[g] != 0) {
                    solved = true;
                    break;
                }
            }
            if (inGroup == 0) isl.positionsSolved = solved ? 1u : 0u;
            for (u32 t = inGroup; t < isl.bodyCount; t += 8) {
                const scalar* r = slab + t * kPoseWords;
                PosBody& p = gpos[t];
                p.x = Vec3(r[0], r[1], r[2]);
                p.q = Quat(Vec3(r[3], r[4], r[5]), r[6]);
                if ((__builtin_bit_cast(u32, r[7]) & kPoseFromGlobal) == 0) p.I = RotateToFrame(bodies[ib[t]].invI, p.q);
            }
        }
    }
}
