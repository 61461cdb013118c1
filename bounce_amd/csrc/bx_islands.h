// bx_islands.h -- island construction and the island solver kernels (included by bx_step.hip).
//
// k_islands: one workgroup per world. The world's body flags, fixture lists, contact endpoints and
// the per-fixture contact edge lists are staged in LDS (or in a global scratch area for worlds too
// large for LDS), then one lane runs the reference's depth-first island search
// (world.cpp:282-420) over them: seeds in body-list order (newest body first), for each popped
// body its fixtures in list order, for each fixture its contacts newest first (prepend order of
// contact_manager.cpp:118-139), then its joints in list order. The output is the island list of
// the world with island-local body indices for every island contact.
//
// k_solve: one lane per island. Gauss-Seidel in the reference's constraint order is a serial
// dependency chain inside an island (consecutive contacts of the DFS order share bodies), so the
// only parallelism is across islands; packing 64 islands into a wavefront keeps every SIMD lane
// of an issued instruction useful.
#pragma once
// (included inside namespace bx by bx_step.hip)


struct StepParams {
    scalar dt, inv_dt;
    u32 velocityIterations, positionIterations;
};

// Builds, for every fixture, its contact edge list in the reference's order (newest contact
// first) as CSR over order-list positions, in global memory (used when pairs are searched before
// the island builder ran in this step).
__device__ void BuildAdjacency(const BatchView& B, const WorldInfo& wi, const WorldState& ws, u32* off, u32* cursor) {
    const u32 nf = wi.fixtureCount;
    const u32 n = ws.contactCount;
    const u32* order = B.contactOrder + wi.contactBase;
    for (u32 f = threadIdx.x; f <= nf; f += blockDim.x) off[f] = 0;
    __syncthreads();
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        const Contact& c = B.contacts[wi.contactBase + order[i]];
        atomicAdd(&off[c.fixtureA], 1u);
        atomicAdd(&off[c.fixtureB], 1u);
    }
    __syncthreads();
    u32 carry = 0;
    for (u32 base = 0; base <= nf; base += blockDim.x) {
        u32 f = base + threadIdx.x;
        u32 v = f <= nf ? off[f] : 0u;
        u32 total;
        u32 p = BlockExclusiveScan(v, &total);
        if (f <= nf) { off[f] = carry + p; cursor[f] = carry + p; }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        u32* edges = B.edgeSlots + 2 * wi.contactBase;
        for (u32 k = n; k-- > 0;) {
            const Contact& c = B.contacts[wi.contactBase + order[k]];
            edges[cursor[c.fixtureA]++] = k;
            edges[cursor[c.fixtureB]++] = k;
        }
    }
    __syncthreads();
}

// The island builder of one world over its working arrays at `base`: LDS (kLds, DS instructions)
// when the world's arrays fit the block's LDS at the current contact count, else its global area.
template <bool kLds, bool kMesh>
__device__ __forceinline__ void BuildIslands(const BatchView& B, u32 w, const WorldInfo& wi, WorldState& ws,
                                             u32* base) {
    const u32 nb = wi.bodyCount, nf = wi.fixtureCount, n = ws.contactCount;
    u32* bflags = base;                 // flags | type << 8
    u32* bfix = bflags + nb;            // fixtureBegin | fixtureCount << 16
    u32* bjoint = bfix + nb;            // jointBegin | jointCount << 16
    u32* bidx = bjoint + nb;            // island index
    u32* stack = bidx + nb;
    u32* flist = stack + nb;            // fixture lists (list order)
    u32* fbody = flist + nf;            // body | sensor << 31
    u32* off = fbody + nf;              // nf + 1
    u32* cursor = off + nf + 1;         // nf + 1
    u32* cfa = cursor + nf + 1;         // by order position, n entries each
    u32* cfb = cfa + n;
    u32* cfl = cfb + n;                 // flags (bit 1 touching) | manifold count << 2
    u32* edges = cfl + n;               // 2 n
    const u32* order = B.contactOrder + wi.contactBase;

    // 1. stage (island flags cleared: world.cpp:265-277)
    for (u32 i = threadIdx.x; i < nb; i += blockDim.x) {
        const Body& b = B.bodies[wi.bodyBase + i];
        bflags[i] = (b.flags & ~kIsland) | (b.type << 8);
        bfix[i] = b.fixtureBegin | (b.fixtureCount << 16);
        bjoint[i] = b.jointBegin | (b.jointCount << 16);
        bidx[i] = b.islandIndex;
    }
    for (u32 f = threadIdx.x; f < nf; f += blockDim.x) {
        flist[f] = B.fixtureList[wi.fixtureListBase + f];
        const Fixture& fx = B.fixtures[wi.fixtureBase + f];
        fbody[f] = fx.body | (fx.isSensor ? 0x80000000u : 0u);
    }
    for (u32 f = threadIdx.x; f <= nf; f += blockDim.x) off[f] = 0;
    for (u32 i = threadIdx.x; i < wi.jointCount; i += blockDim.x) B.joints[wi.jointBase + i].islandFlag = 0;
    __syncthreads();
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        const Contact& c = B.contacts[wi.contactBase + order[i]];
        cfa[i] = c.fixtureA;
        cfb[i] = c.fixtureB;
        cfl[i] = (c.flags & 2u) | (Min(c.manifoldCount, kMaxManifolds) << 2);
        atomicAdd(&off[c.fixtureA], 1u);
        atomicAdd(&off[c.fixtureB], 1u);
    }
    __syncthreads();
    // 2. adjacency: CSR offsets, then edges newest contact first
    u32 carry = 0;
    for (u32 b0 = 0; b0 <= nf; b0 += blockDim.x) {
        u32 f = b0 + threadIdx.x;
        u32 v = f <= nf ? off[f] : 0u;
        u32 total;
        u32 p = BlockExclusiveScan(v, &total);
        if (f <= nf) { off[f] = carry + p; cursor[f] = carry + p; }
        carry += total;
        __syncthreads();
    }
    // every fixture's contacts, newest (highest order position) first: filled in any order, then
    // each contact placed at its rank among its fixture's contacts in the pair search's copy
    // (B.edgeSlots), which the island search reads back
    for (u32 k = threadIdx.x; k < n; k += blockDim.x) {
        edges[atomicAdd(&cursor[cfa[k]], 1u)] = k;
        edges[atomicAdd(&cursor[cfb[k]], 1u)] = k;
    }
    __syncthreads();
    u32* slots = B.edgeSlots + 2 * wi.contactBase;
    for (u32 k = threadIdx.x; k < n; k += blockDim.x) {
        const u32 ends[2] = {cfa[k], cfb[k]};
        for (u32 s = 0; s < 2; ++s) {
            const u32 f = ends[s];
            u32 rank = 0;
            for (u32 e = off[f]; e < off[f + 1]; ++e) rank += edges[e] > k ? 1u : 0u;
            slots[off[f] + rank] = k;
        }
    }
    for (u32 f = threadIdx.x; f <= nf; f += blockDim.x) B.edgeOffset[wi.edgeBase + f] = off[f];
    __syncthreads();
    for (u32 e = threadIdx.x; e < 2 * n; e += blockDim.x) edges[e] = slots[e];
    __syncthreads();

    // 3. depth-first search (one lane)
    if (threadIdx.x == 0) {
        IslandInfo* islands = B.islands + wi.bodyBase;
        u32* ibodies = B.islandBodies + wi.islandBodyBase;
        uint4* icontacts = B.islandContacts + wi.conBase;
        uint4* ijoints = B.islandJoints + wi.jointBase;
        u32 islandCount = 0, nbo = 0, nc = 0, nj = 0, overflow = 0;
        for (u32 s = nb; s-- > 0;) {
            u32 sf = bflags[s];
            if (sf & kIsland) continue;
            if ((sf & kAwake) == 0) continue;
            if ((sf >> 8) == kStatic) continue;
            IslandInfo info;
            info.bodyStart = nbo;
            info.contactStart = nc;
            info.jointStart = nj;
            u32 bodyCount = 0, stackCount = 0;
            stack[stackCount++] = s;
            bflags[s] = sf | kIsland;
            while (stackCount > 0) {
                u32 bi = stack[--stackCount];
                bidx[bi] = bodyCount;
                if (nbo < wi.islandBodyCap) ibodies[nbo] = bi;
                else overflow = 1;
                ++nbo;
                ++bodyCount;
                u32 bf = bflags[bi];
                if ((bf >> 8) == kStatic) continue;
                bflags[bi] = bf | kAwake;
                u32 fb = bfix[bi] & 0xFFFFu, fc = bfix[bi] >> 16;
                for (u32 fi = 0; fi < fc; ++fi) {
                    u32 f = flist[fb + fi];
                    for (u32 e = off[f]; e < off[f + 1]; ++e) {
                        u32 k = edges[e];
                        u32 cf = cfl[k];
                        if (cf & 1u) continue;
                        if ((cf & 2u) == 0) continue;
                        u32 fa = cfa[k], fbb = cfb[k];
                        if ((fbody[fa] | fbody[fbb]) & 0x80000000u) continue;  // sensors
#if BX_DEVICE_CALLBACKS
                        // b3ContactFilter::ShouldRespond (world.cpp:351-357)
                        if (wi.filter && !FilterBit(wi.filter + FilterWords(nf), nf, fa, fbb)) continue;
#endif
                        // one constraint entry per manifold (a mesh contact has up to three,
                        // solved one after the other as b3ContactSolver does; a touching convex
                        // contact has one)
                        // (the order position k becomes the contact slot in the fix-up below, which
                        // reads the order list in batches instead of once per edge here)
                        if (kMesh) {
                            for (u32 j = 0; j < (cf >> 2); ++j)
                                icontacts[nc++] = make_uint4(k | (j << 30), fbody[fa], fbody[fbb], 0u);
                        } else {
                            icontacts[nc++] = make_uint4(k, fbody[fa], fbody[fbb], 0u);
                        }
                        cfl[k] = cf | 1u;
                        u32 other = (fa == f ? fbody[fbb] : fbody[fa]) & 0x7FFFFFFFu;
                        if (bflags[other] & kIsland) continue;
                        stack[stackCount++] = other;
                        bflags[other] |= kIsland;
                    }
                }
                u32 jb = bjoint[bi] & 0xFFFFu, jc = bjoint[bi] >> 16;
                for (u32 ji = 0; ji < jc; ++ji) {
                    JointEdge je = B.jointEdges[wi.jointEdgeBase + jb + ji];
                    Joint& j = B.joints[wi.jointBase + je.joint];
                    if (j.islandFlag) continue;
                    ijoints[nj++] = make_uint4(je.joint, j.bodyA, j.bodyB, 0u);
                    j.islandFlag = 1;
                    if (bflags[je.other] & kIsland) continue;
                    stack[stackCount++] = je.other;
                    bflags[je.other] |= kIsland;
                }
            }
            info.bodyCount = bodyCount;
            info.contactCount = nc - info.contactStart;
            info.jointCount = nj - info.jointStart;
            // island-local indices of each contact's bodies (a static body keeps only its index in
            // the island being built, exactly what b3ContactSolver reads for this island)
            // (eight entries per batch: their global loads are issued together)
            auto fixContact = [&](uint4& e, u32 slot) {
                e.x = slot | (e.x & 0xC0000000u);
                e.y = bidx[e.y & 0x7FFFFFFFu];
                e.z = bidx[e.z & 0x7FFFFFFFu];
                e.w = info.bodyStart;
            };
            u32 k = info.contactStart;
            for (; k + 8 <= nc; k += 8) {
                uint4 e[8];
                u32 slot[8];
#pragma unroll
                for (u32 t = 0; t < 8; ++t) e[t] = icontacts[k + t];
#pragma unroll
                for (u32 t = 0; t < 8; ++t) slot[t] = order[e[t].x & 0x3FFFFFFFu];
#pragma unroll
                for (u32 t = 0; t < 8; ++t) {
                    fixContact(e[t], slot[t]);
                    icontacts[k + t] = e[t];
                }
            }
            for (; k < nc; ++k) {
                uint4 e = icontacts[k];
                fixContact(e, order[e.x & 0x3FFFFFFFu]);
                icontacts[k] = e;
            }
            for (u32 k = info.jointStart; k < nj; ++k) {  // same for the island's joints
                uint4 e = ijoints[k];
                e.y = bidx[e.y];
                e.z = bidx[e.z];
                e.w = info.bodyStart;
                ijoints[k] = e;
            }
            islands[islandCount++] = info;
            for (u32 k0 = 0; k0 < bodyCount; k0 += 8) {  // world.cpp:411-419, eight loads per batch
                u32 bis[8];
#pragma unroll
                for (u32 t = 0; t < 8; ++t) {
                    const u32 at = info.bodyStart + k0 + t;
                    bis[t] = k0 + t < bodyCount && at < wi.islandBodyCap ? ibodies[at] : kNull;
                }
#pragma unroll
                for (u32 t = 0; t < 8; ++t)
                    if (bis[t] != kNull && (bflags[bis[t]] >> 8) == kStatic) bflags[bis[t]] &= ~kIsland;
            }
        }
        ws.islandCount = islandCount;
        ws.islandBodyCount = nbo;
        ws.islandContactCount = nc;
        ws.islandJointCount = nj;
        if (overflow) ws.status |= 0x2u;
        // publish the world's islands to the solver's work list
        u32 entries = wi.serialIslands ? (islandCount ? 1u : 0u) : islandCount;
        u32 at = atomicAdd(B.islandListCount, entries);
        if (wi.serialIslands) {
            if (entries) B.islandList[at] = make_uint2(w, kNull);
        } else {
            for (u32 i = 0; i < islandCount; ++i) B.islandList[at + i] = make_uint2(w, i);
        }
    }
    __syncthreads();
    // 4. write back body flags / island indices
    for (u32 i = threadIdx.x; i < nb; i += blockDim.x) {
        Body& b = B.bodies[wi.bodyBase + i];
        b.flags = bflags[i] & 0xFFu;
        b.islandIndex = bidx[i];
    }
}

template <bool kMesh>
__global__ void __launch_bounds__(kBlock) k_islands(BatchView B) {
    extern __shared__ u32 lds[];
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    WorldState& ws = B.wstate[w];
    if (IslandScratchWords(wi.bodyCount, wi.fixtureCount, ws.contactCount) <= B.islandLdsWords)
        BuildIslands<true, kMesh>(B, w, wi, ws, lds);
    else
        BuildIslands<false, kMesh>(B, w, wi, ws, B.islandScratch + wi.islandScratchBase);
}

