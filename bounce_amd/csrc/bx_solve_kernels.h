// bx_solve_kernels.h -- b3Island::Solve (island.cpp:81-307) as a sequence of kernels
// (included inside namespace bx by bx_step.hip).
//
// Only the Gauss-Seidel sweeps are sequential in the reference's sense: warm starting, the
// velocity iterations and the position iterations visit the island's contacts in island order
// and each visit reads the bodies the previous visits wrote. Everything else in b3Island::Solve
// is independent per body or per contact (velocity integration, constraint set-up, storing the
// impulses, position integration, copy-back), so it runs as fully parallel kernels:
//
//   k_integrate_vel   per island body    island.cpp:86-135
//   k_init_contacts   per island contact contact_solver.cpp:28-270
//   k_solve_velocity  per island (lane)  warm start + velocity iterations (island.cpp:154-181)
//   k_mid             per contact/body   StoreImpulses + position integration (island.cpp:183-229)
//   k_solve_position  per island (lane)  position iterations (island.cpp:231-248)
//   k_finish          per body/island    copy back + sleep (island.cpp:249-306)

#ifndef BX_TWO_BUFFERS
#define BX_TWO_BUFFERS 1
#endif
#ifndef BX_POS_LANES
#define BX_POS_LANES 64
#endif
#ifndef BX_POS_TWO_BUFFERS
#define BX_POS_TWO_BUFFERS 0
#endif
constexpr u32 kPosLanes = BX_POS_LANES;      // islands per 64-wide workgroup in the position sweeps
#ifndef BX_POS_SPREAD_LDS
#define BX_POS_SPREAD_LDS 0
#endif
// LDS requested (and not used) by the position-sweep workgroups so that at most one runs per CU:
// the sweep of a wavefront is a serial chain whose time grows with every other wavefront sharing
// its CU, so the islands are spread over all CUs instead.
constexpr u32 kPosSpreadLds = BX_POS_SPREAD_LDS;

__global__ void __launch_bounds__(kBlock) k_integrate_vel(BatchView B, StepParams sp) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const WorldState& ws = B.wstate[w];
    const u32 n = Min(ws.islandBodyCount, wi.islandBodyCap);
    Body* bodies = B.bodies + wi.bodyBase;
    const u32* ib = B.islandBodies + wi.islandBodyBase;
    VelBody* vel = B.velBodies + wi.islandBodyBase;
    PosBody* pos = B.posBodies + wi.islandBodyBase;
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        Body& b = bodies[ib[i]];
        IntegrateVelocity(b, wi.gravity, sp.dt, vel[i], pos[i]);
        b.x0 = b.x;  // static bodies of several islands write the same values
        b.q0 = b.q;
    }
}

__global__ void __launch_bounds__(kBlock) k_init_contacts(BatchView B, StepParams sp) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const WorldState& ws = B.wstate[w];
    const scalar dtRatio = ws.inv_dt0 * sp.dt;
    const bool warm = (wi.flags & 0x2u) == 0;
    const uint4* ic = B.islandContacts + wi.conBase;
    const Body* bodies = B.bodies + wi.bodyBase;
    for (u32 k = threadIdx.x; k < ws.islandContactCount; k += blockDim.x) {
        const uint4 e = ic[k];
        const u32 slot = e.x & 0x3FFFFFFFu;
        Contact& c = B.contacts[wi.contactBase + slot];
        const Fixture& fA = B.fixtures[wi.fixtureBase + c.fixtureA];
        const Fixture& fB = B.fixtures[wi.fixtureBase + c.fixtureB];
        const VelBody* vel = B.velBodies + wi.islandBodyBase + e.w;
        const PosBody* pos = B.posBodies + wi.islandBodyBase + e.w;
        VelCon vc;
        PosCon pc;
        InitContact(vc, pc, ContactManifold(B, wi, c, slot, e.x >> 30), c.manifoldCount, fA, fB, bodies[fA.body],
                    bodies[fB.body], e.y, e.z, vel, pos, dtRatio, warm);
        B.velCons[wi.conBase + k] = vc;
        B.posCons[wi.conBase + k] = pc;
    }
    // b3JointSolver::InitializeVelocityConstraints: independent per joint (revolute_joint.cpp:78-231)
    const uint4* ij = B.islandJoints + wi.jointBase;
    for (u32 k = threadIdx.x; k < ws.islandJointCount; k += blockDim.x) {
        const uint4 e = ij[k];
        Joint& j = B.joints[wi.jointBase + e.x];
        JointCon jc;
        InitJointCon(jc, j, bodies[j.bodyA], bodies[j.bodyB], e.y, e.z, B.posBodies + wi.islandBodyBase + e.w,
                     sp.dt, sp.inv_dt, dtRatio, warm);
        B.jointCons[wi.jointBase + k] = jc;
    }
}

// Warm start and velocity iterations of one island on one lane. kInLds selects at compile time
// whether the velocities live in the lane's LDS slab (LDS instructions, which do not share the
// in-order vector-memory counter with the constraint prefetch) or stay in global memory.
// One contact of a velocity sweep on body velocities `vel`: read the two bodies, request the
// constraint `ahead` (the next one, clamped to the last), apply `op`, write the bodies back.
// Two register buffers alternate between the current and the requested constraint, so the
// prefetched constraint is never copied between registers.
template <class V, class Op>
__device__ __forceinline__ void VelocityStep(VelCon& cur, VelCon& ahead, const VelCon* src, V* vel, Op op) {
    Vec3 vA = vel[cur.indexA].v, wA = vel[cur.indexA].w, vB = vel[cur.indexB].v, wB = vel[cur.indexB].w;
    ahead = *src;
    op(cur, vA, wA, vB, wB);
    vel[cur.indexA].v = vA;
    vel[cur.indexA].w = wA;
    vel[cur.indexB].v = vB;
    vel[cur.indexB].w = wB;
}

// A sweep over the island's contacts in order. kTwoBuffers unrolls by two so that the two
// register buffers swap roles instead of the prefetched constraint being copied each contact
// (twice the code; used for the warm start and the iteration sweeps on LDS velocities).
template <bool kTwoBuffers, class V, class Op>
__device__ __forceinline__ void ContactSweep(VelCon* vcs, u32 nc, V* vel, Op op) {
    VelCon a = vcs[0], b;
    if (!kTwoBuffers) {
        for (u32 k = 0; k < nc; ++k) {
            VelocityStep(a, b, vcs + Min(k + 1, nc - 1), vel, op);
            op.after(a, vcs[k]);
            a = b;
        }
        return;
    }
    for (u32 k = 0;;) {
        VelocityStep(a, b, vcs + Min(k + 1, nc - 1), vel, op);
        op.after(a, vcs[k]);
        if (++k == nc) break;
        VelocityStep(b, a, vcs + Min(k + 1, nc - 1), vel, op);
        op.after(b, vcs[k]);
        if (++k == nc) break;
    }
}

struct WarmStartOp {
    __device__ __forceinline__ void operator()(VelCon& vc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) const {
        WarmStartContact(vc, vA, wA, vB, wB);
    }
    __device__ __forceinline__ void after(const VelCon&, VelCon&) const {}
};

struct SolveOp {
    __device__ __forceinline__ void operator()(VelCon& vc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) const {
        SolveContactVelocity(vc, vA, wA, vB, wB);
    }
    // the impulses accumulated by this sweep (the next sweep and StoreImpulses read them)
    __device__ __forceinline__ void after(const VelCon& vc, VelCon& dst) const {
        dst.tangentImpulse = vc.tangentImpulse;
        dst.motorImpulse = vc.motorImpulse;
#pragma unroll
        for (u32 p = 0; p < kMaxManifoldPoints; ++p)
            if (p < vc.pointCount) dst.points[p].normalImpulse = vc.points[p].normalImpulse;
    }
};

// Warm start and velocity iterations of one island on one lane. kInLds selects at compile time
// whether the velocities live in the lane's LDS slab (LDS instructions, which do not share the
// in-order vector-memory counter with the constraint prefetch) or stay in global memory.
template <bool kInLds>
__device__ __forceinline__ void VelocitySweeps(VelCon* vcs, u32 nc, JointCon* jcs, u32 nj, VelBody* vel,
                                               u32 iterations, bool warm, scalar dt) {
    // Each contact's body velocities are read first and only then the next contact's constraint is
    // requested (clamped index, no branch), so waiting for the bodies never waits for the prefetch.
    // Order of island.cpp:154-181: contact warm start, joint warm start, then per iteration all
    // joints followed by all contacts.
    if (warm && nc > 0) ContactSweep<kInLds && BX_TWO_BUFFERS>(vcs, nc, vel, WarmStartOp());
    if (warm)
        for (u32 k = 0; k < nj; ++k) WarmStartJoint(jcs[k], vel);
    for (u32 it = 0; it < iterations && (nc > 0 || nj > 0); ++it) {
        for (u32 k = 0; k < nj; ++k) SolveJointVelocity(jcs[k], vel, dt);
        if (nc > 0) ContactSweep<kInLds && BX_TWO_BUFFERS>(vcs, nc, vel, SolveOp());
    }
}

// Bodies of the largest island of an island-list entry (islands first..last of one world).
__device__ __forceinline__ u32 EntryLargestIsland(const BatchView& B, const WorldInfo& wi, u32 first, u32 last) {
    u32 most = 0;
    for (u32 k = first; k < last; ++k) most = Max(most, B.islands[wi.bodyBase + k].bodyCount);
    return most;
}

// kLanes islands per workgroup of one wavefront (its first kLanes lanes), each island's velocities
// in an LDS slab of kSlab bodies when they fit. Only entries whose largest island has more than
// minBodies bodies are solved (0: every entry; the eight-lane layouts leave larger islands here).
template <u32 kLanes, u32 kSlab>
__global__ void __launch_bounds__(64) k_solve_velocity(BatchView B, StepParams sp, u32 minBodies) {
    static_assert(kLanes <= 64 && kLanes * kSlab * sizeof(VelBody) <= 96 * 1024, "velocity sweep layout");
    __shared__ VelBody slabs[kLanes][kSlab];
    const u32 slot = threadIdx.x;
    if (slot >= kLanes) return;
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * kLanes + slot; i < count; i += gridDim.x * kLanes) {
        const uint2 e = B.islandList[i];
        BX_WORLD_HOOK(B, e.x);
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const bool warm = (wi.flags & 0x2u) == 0;
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        if (minBodies && EntryLargestIsland(B, wi, first, last) <= minBodies) continue;
        for (u32 k = first; k < last; ++k) {
            const IslandInfo isl = B.islands[wi.bodyBase + k];
            VelBody* gvel = B.velBodies + wi.islandBodyBase + isl.bodyStart;
            VelCon* vcs = B.velCons + wi.conBase + isl.contactStart;
            JointCon* jcs = B.jointCons + wi.jointBase + isl.jointStart;
            if (isl.bodyCount <= kSlab) {
                VelBody* slab = slabs[slot];
                for (u32 b = 0; b < isl.bodyCount; ++b) slab[b] = gvel[b];
                VelocitySweeps<true>(vcs, isl.contactCount, jcs, isl.jointCount, slab, sp.velocityIterations, warm,
                                     sp.dt);
                for (u32 b = 0; b < isl.bodyCount; ++b) gvel[b] = slab[b];
            } else {
                VelocitySweeps<false>(vcs, isl.contactCount, jcs, isl.jointCount, gvel, sp.velocityIterations, warm,
                                      sp.dt);
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_mid(BatchView B, StepParams sp) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const WorldState& ws = B.wstate[w];
    if ((wi.flags & 0x2u) == 0) {  // b3ContactSolver::StoreImpulses (contact_solver.cpp:488-515)
        const uint4* ic = B.islandContacts + wi.conBase;
        for (u32 k = threadIdx.x; k < ws.islandContactCount; k += blockDim.x) {
            const u32 slot = ic[k].x & 0x3FFFFFFFu;
            Contact& c = B.contacts[wi.contactBase + slot];
            if (c.manifoldCount == 0) continue;
            Manifold& m = ContactManifold(B, wi, c, slot, ic[k].x >> 30);
            const VelCon& vc = B.velCons[wi.conBase + k];
            m.tangentImpulse = vc.tangentImpulse;
            m.motorImpulse = vc.motorImpulse;
            for (u32 p = 0; p < m.pointCount; ++p) m.points[p].normalImpulse = vc.points[p].normalImpulse;
        }
    }
    // the joints keep their impulses for the next step's warm start (revolute_joint.h members)
    const uint4* ij = B.islandJoints + wi.jointBase;
    for (u32 k = threadIdx.x; k < ws.islandJointCount; k += blockDim.x) {
        Joint& j = B.joints[wi.jointBase + ij[k].x];
        const JointCon& jc = B.jointCons[wi.jointBase + k];
        j.motorImpulse = jc.motorImpulse;
        j.limitImpulse = jc.limitImpulse;
        j.linearImpulse = jc.linearImpulse;
        j.angularImpulse = jc.angularImpulse;
        j.angularImpulse3 = jc.angularImpulse3;
        j.impulse1 = jc.impulse1;
        j.linearImpulse2 = jc.linearImpulse2;
        j.springImpulse = jc.springImpulse;
    }
    const u32 n = Min(ws.islandBodyCount, wi.islandBodyCap);
    const u32* ib = B.islandBodies + wi.islandBodyBase;
    for (u32 i = threadIdx.x; i < n; i += blockDim.x)
        IntegratePosition(B.velBodies[wi.islandBodyBase + i], B.posBodies[wi.islandBodyBase + i],
                          B.bodies[wi.bodyBase + ib[i]], sp.dt);
}

// One contact of a position sweep: read the two bodies, request the constraint `ahead`, solve.
__device__ __forceinline__ scalar PositionStep(const PosCon& cur, PosCon& ahead, const PosCon* src, PosBody* pos,
                                               scalar minSeparation) {
    PosBody A = pos[cur.indexA];
    PosBody Bb = pos[cur.indexB];
    ahead = *src;
    minSeparation = SolveContactPosition(cur, A.x, A.q, A.I, Bb.x, Bb.q, Bb.I, minSeparation);
    pos[cur.indexA] = A;
    pos[cur.indexB] = Bb;
    return minSeparation;
}

__device__ __forceinline__ bool PositionSweeps(const PosCon* pcs, u32 nc, const JointCon* jcs, u32 nj, PosBody* pos,
                                               u32 iterations) {
    for (u32 it = 0; it < iterations; ++it) {
        scalar minSeparation = 0;
        if (nc > 0 && BX_POS_TWO_BUFFERS) {  // two register buffers swap roles (no copy of the prefetch)
            PosCon a = pcs[0], b;
            for (u32 k = 0;;) {
                minSeparation = PositionStep(a, b, pcs + Min(k + 1, nc - 1), pos, minSeparation);
                if (++k == nc) break;
                minSeparation = PositionStep(b, a, pcs + Min(k + 1, nc - 1), pos, minSeparation);
                if (++k == nc) break;
            }
        } else if (nc > 0) {
            PosCon a = pcs[0], b;
            for (u32 k = 0; k < nc; ++k) {
                minSeparation = PositionStep(a, b, pcs + Min(k + 1, nc - 1), pos, minSeparation);
                a = b;
            }
        }
        bool contactsSolved = minSeparation >= scalar(-3) * kLinearSlop;
        bool jointsSolved = true;  // joint_solver.cpp:55-64
        for (u32 k = 0; k < nj; ++k) {
            bool solved = SolveJointPosition(jcs[k], pos);
            jointsSolved = jointsSolved && solved;
        }
        if (contactsSolved && jointsSolved) return true;
    }
    return false;
}

// A lane per island; only entries whose largest island has more than minBodies bodies (0: every
// entry; the eight-lane layouts leave larger islands here).
__global__ void __launch_bounds__(64) k_solve_position(BatchView B, StepParams sp, u32 minBodies) {
    if (threadIdx.x >= kPosLanes) return;
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * kPosLanes + threadIdx.x; i < count; i += gridDim.x * kPosLanes) {
        const uint2 e = B.islandList[i];
        BX_WORLD_HOOK(B, e.x);
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        if (minBodies && EntryLargestIsland(B, wi, first, last) <= minBodies) continue;
        for (u32 k = first; k < last; ++k) {
            IslandInfo& isl = B.islands[wi.bodyBase + k];
            isl.positionsSolved = PositionSweeps(B.posCons + wi.conBase + isl.contactStart, isl.contactCount,
                                                 B.jointCons + wi.jointBase + isl.jointStart, isl.jointCount,
                                                 B.posBodies + wi.islandBodyBase + isl.bodyStart,
                                                 sp.positionIterations) ? 1u : 0u;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_finish(BatchView B, StepParams sp) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const WorldState& ws = B.wstate[w];
    const u32 n = Min(ws.islandBodyCount, wi.islandBodyCap);
    const u32* ib = B.islandBodies + wi.islandBodyBase;
    Body* bodies = B.bodies + wi.bodyBase;
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {  // island.cpp:249-261
        Body& b = bodies[ib[i]];
        const PosBody p = B.posBodies[wi.islandBodyBase + i];
        const VelBody v = B.velBodies[wi.islandBodyBase + i];
        b.x = p.x;
        b.q = p.q;
        b.worldInvI = p.I;
        b.v = v.v;
        b.w = v.w;
        b.xfT = b.x - Mul(b.q, b.localCenter);
    }
    if ((wi.flags & 0x1u) == 0) return;
    __syncthreads();
    // island.cpp:266-306: one thread per island (bodies of different islands are disjoint except
    // static bodies, which the timer skips)
    const scalar h = sp.dt;
    for (u32 k = threadIdx.x; k < ws.islandCount; k += blockDim.x) {
        const IslandInfo isl = B.islands[wi.bodyBase + k];
        scalar minSleepTime = kMaxScalar;
        for (u32 i = 0; i < isl.bodyCount; ++i) {
            Body& b = bodies[ib[isl.bodyStart + i]];
            if (b.type == kStatic) continue;
            scalar sqrLin = Dot(b.v, b.v);
            scalar sqrAng = Dot(b.w, b.w);
            if ((b.flags & kAutoSleep) == 0 || sqrLin > b.linTol * b.linTol || sqrAng > b.angTol * b.angTol)
                b.sleepTime = 0;
            else
                b.sleepTime += h;
            minSleepTime = Min(minSleepTime, b.sleepTime);
        }
        if (minSleepTime >= kTimeToSleep && isl.positionsSolved) {
            for (u32 i = 0; i < isl.bodyCount; ++i) {
                Body& b = bodies[ib[isl.bodyStart + i]];
                atomicAnd(&b.flags, ~kAwake);
                b.sleepTime = 0;
                b.force.SetZero();
                b.torque.SetZero();
                b.v.SetZero();
                b.w.SetZero();
            }
        }
    }
}
