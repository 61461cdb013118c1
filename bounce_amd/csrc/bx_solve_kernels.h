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

constexpr u32 kSolveLanes = 8;     // islands per 64-wide workgroup in the sweep kernels
constexpr u32 kSlabBodies = 256;   // velocities of islands up to this size live in LDS

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
    const uint4* ic = B.islandContacts + wi.contactBase;
    const Body* bodies = B.bodies + wi.bodyBase;
    for (u32 k = threadIdx.x; k < ws.islandContactCount; k += blockDim.x) {
        const uint4 e = ic[k];
        const Contact& c = B.contacts[wi.contactBase + e.x];
        const Fixture& fA = B.fixtures[wi.fixtureBase + c.fixtureA];
        const Fixture& fB = B.fixtures[wi.fixtureBase + c.fixtureB];
        const VelBody* vel = B.velBodies + wi.islandBodyBase + e.w;
        const PosBody* pos = B.posBodies + wi.islandBodyBase + e.w;
        VelCon vc;
        PosCon pc;
        InitContact(vc, pc, c, fA, fB, bodies[fA.body], bodies[fB.body], e.y, e.z, vel, pos, dtRatio, warm);
        B.velCons[wi.contactBase + k] = vc;
        B.posCons[wi.contactBase + k] = pc;
    }
    // b3JointSolver::InitializeVelocityConstraints: independent per joint (revolute_joint.cpp:78-231)
    const uint4* ij = B.islandJoints + wi.jointBase;
    for (u32 k = threadIdx.x; k < ws.islandJointCount; k += blockDim.x) {
        const uint4 e = ij[k];
        Joint& j = B.joints[wi.jointBase + e.x];
        JointCon jc;
        InitJointCon(jc, j, bodies[j.bodyA], bodies[j.bodyB], e.y, e.z, B.posBodies + wi.islandBodyBase + e.w,
                     dtRatio, warm);
        B.jointCons[wi.jointBase + k] = jc;
    }
}

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
    if (warm && nc > 0) {
        VelCon next = vcs[0];
        for (u32 k = 0; k < nc; ++k) {
            const VelCon vc = next;
            Vec3 vA = vel[vc.indexA].v, wA = vel[vc.indexA].w, vB = vel[vc.indexB].v, wB = vel[vc.indexB].w;
            next = vcs[Min(k + 1, nc - 1)];
            WarmStartContact(vc, vA, wA, vB, wB);
            vel[vc.indexA].v = vA; vel[vc.indexA].w = wA;
            vel[vc.indexB].v = vB; vel[vc.indexB].w = wB;
        }
    }
    if (warm)
        for (u32 k = 0; k < nj; ++k) WarmStartJoint(jcs[k], vel);
    for (u32 it = 0; it < iterations && (nc > 0 || nj > 0); ++it) {
        for (u32 k = 0; k < nj; ++k) SolveJointVelocity(jcs[k], vel, dt);
        if (nc == 0) continue;
        VelCon next = vcs[0];
        for (u32 k = 0; k < nc; ++k) {
            VelCon vc = next;
            Vec3 vA = vel[vc.indexA].v, wA = vel[vc.indexA].w, vB = vel[vc.indexB].v, wB = vel[vc.indexB].w;
            next = vcs[Min(k + 1, nc - 1)];
            SolveContactVelocity(vc, vA, wA, vB, wB);
            vel[vc.indexA].v = vA; vel[vc.indexA].w = wA;
            vel[vc.indexB].v = vB; vel[vc.indexB].w = wB;
            VelCon& dst = vcs[k];
            dst.tangentImpulse = vc.tangentImpulse;
            dst.motorImpulse = vc.motorImpulse;
#pragma unroll
            for (u32 p = 0; p < kMaxManifoldPoints; ++p)
                if (p < vc.pointCount) dst.points[p].normalImpulse = vc.points[p].normalImpulse;
        }
    }
}

__global__ void __launch_bounds__(64) k_solve_velocity(BatchView B, StepParams sp) {
    __shared__ VelBody slabs[kSolveLanes][kSlabBodies];
    if (threadIdx.x >= kSolveLanes) return;
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * kSolveLanes + threadIdx.x; i < count; i += gridDim.x * kSolveLanes) {
        const uint2 e = B.islandList[i];
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const bool warm = (wi.flags & 0x2u) == 0;
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        for (u32 k = first; k < last; ++k) {
            const IslandInfo isl = B.islands[wi.bodyBase + k];
            VelBody* gvel = B.velBodies + wi.islandBodyBase + isl.bodyStart;
            VelCon* vcs = B.velCons + wi.contactBase + isl.contactStart;
            JointCon* jcs = B.jointCons + wi.jointBase + isl.jointStart;
            if (isl.bodyCount <= kSlabBodies) {
                VelBody* slab = slabs[threadIdx.x];
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
        const uint4* ic = B.islandContacts + wi.contactBase;
        for (u32 k = threadIdx.x; k < ws.islandContactCount; k += blockDim.x) {
            Contact& c = B.contacts[wi.contactBase + ic[k].x];
            if (c.manifoldCount == 0) continue;
            const VelCon& vc = B.velCons[wi.contactBase + k];
            c.manifold.tangentImpulse = vc.tangentImpulse;
            c.manifold.motorImpulse = vc.motorImpulse;
            for (u32 p = 0; p < c.manifold.pointCount; ++p) c.manifold.points[p].normalImpulse = vc.points[p].normalImpulse;
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
    }
    const u32 n = Min(ws.islandBodyCount, wi.islandBodyCap);
    const u32* ib = B.islandBodies + wi.islandBodyBase;
    for (u32 i = threadIdx.x; i < n; i += blockDim.x)
        IntegratePosition(B.velBodies[wi.islandBodyBase + i], B.posBodies[wi.islandBodyBase + i],
                          B.bodies[wi.bodyBase + ib[i]], sp.dt);
}

__device__ __forceinline__ bool PositionSweeps(const PosCon* pcs, u32 nc, const JointCon* jcs, u32 nj, PosBody* pos,
                                               u32 iterations) {
    for (u32 it = 0; it < iterations; ++it) {
        scalar minSeparation = 0;
        PosCon next;
        if (nc > 0) next = pcs[0];
        for (u32 k = 0; k < nc; ++k) {
            const PosCon pc = next;
            PosBody A = pos[pc.indexA];
            PosBody Bb = pos[pc.indexB];
            next = pcs[Min(k + 1, nc - 1)];
            minSeparation = SolveContactPosition(pc, A.x, A.q, A.I, Bb.x, Bb.q, Bb.I, minSeparation);
            pos[pc.indexA] = A;
            pos[pc.indexB] = Bb;
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

__global__ void __launch_bounds__(64) k_solve_position(BatchView B, StepParams sp) {
    const u32 count = *B.islandListCount;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint2 e = B.islandList[i];
        const WorldInfo wi = B.worlds[e.x];
        const WorldState& ws = B.wstate[e.x];
        const u32 first = e.y == kNull ? 0u : e.y;
        const u32 last = e.y == kNull ? ws.islandCount : e.y + 1;
        for (u32 k = first; k < last; ++k) {
            IslandInfo& isl = B.islands[wi.bodyBase + k];
            isl.positionsSolved = PositionSweeps(B.posCons + wi.contactBase + isl.contactStart, isl.contactCount,
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
