// Glued blocks: the rigid projection (K/RigidBlocks.kt:280-357,
// S/rigid_reduce.comp, S/rigid_apply.comp, S/rigid_remap.comp,
// S/rigid_layout.glsl). A block's members are listed in whole tiles of 256
// slots (padding slots hold -1), so a tile belongs to one block. A projection
// sums each tile's member positions and velocities about the block's
// previous centre, fits each block's rigid motion to the sums (centre,
// best-fit rotation, velocity and spin that keep momentum and angular
// momentum), settles the pair contact memory after a contact pass, and snaps
// every member onto the fitted motion.
//
// Three launches where the reference has two: tile sums, then one workgroup
// per block fits it, then the snap. The reference fits a block in whichever
// of its tiles finishes last; the fit adds the tile sums in tile order either
// way, so the results are the same, and a launch boundary replaces the
// coherent partials and the completion counter (the tiles of a block land on
// different XCDs, each with its own L2).
//
// Float expressions follow the shaders operation for operation (the library
// is built without contraction; division and sqrt round correctly); the
// sums over a tile's slots use the shader's tree (stride 128 down to 1).
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace gs {

namespace {

constexpr int RIGID_TILE = 256;
constexpr float GUARD_SCALE = 65536.0f;
constexpr int IMPACT_SPEED_MASK = 0xffffff;
constexpr int IMPACT_AGE_SHIFT = 24;
constexpr int IMPACT_PUSHING_PASSES = 32;
constexpr int CONTACT_TOUCHING = 1;
constexpr int CONTACT_PUSHING = 2;
constexpr int IMPACT_IDLE_PASSES = 4;
// GlueShape.ROW_PIXELS: rows of member centres a hundredth of a pixel apart
constexpr float HULL_ROW_PIXELS = 0.01f;

struct RigidArgs {
    float width, height;   // viewport, px
    int bounded;
    int contactPass;       // after a contact pass: settle the contact memory, stop against pushes
    int groupCount;        // group indices in use
    float substepDt;       // s
    float bodyRadius;      // px
    float blockFriction;
};

// S/rigid_reduce.comp:60-69: the shader's tree over the 256 slots of a tile.
__device__ __forceinline__ void tree_sum(float4* sa, float4* sb, int lane) {
    __syncthreads();
    for (int stride = RIGID_TILE / 2; stride > 0; stride >>= 1) {
        if (lane < stride) {
            float4 a = sa[lane], a2 = sa[lane + stride];
            float4 b = sb[lane], b2 = sb[lane + stride];
            sa[lane] = make_float4(a.x + a2.x, a.y + a2.y, a.z + a2.z, a.w + a2.w);
            sb[lane] = make_float4(b.x + b2.x, b.y + b2.y, b.z + b2.z, b.w + b2.w);
        }
        __syncthreads();
    }
}

// S/rigid_reduce.comp:203-229: one tile's sums of the member positions about
// the block's previous centre (minimum image on the torus) and velocities,
// the rotation fit terms r0 x p and r0 . p, and the angular momentum term
// (R r0) x v, R the previous rotation.
__global__ void __launch_bounds__(RIGID_TILE) k_rigid_reduce(
    const float4* __restrict__ body, const int* __restrict__ members,
    const float2* __restrict__ rest, const int* __restrict__ tileGroups,
    const RigidGroup* __restrict__ groups, float4* __restrict__ partials, RigidArgs a) {
    __shared__ float4 sa[RIGID_TILE], sb[RIGID_TILE];
    const int tile = blockIdx.x, lane = threadIdx.x;
    // the slot's member and rest offset load with the tile's block
    const int member = tile * RIGID_TILE + lane;
    const int b = members[member];
    const float2 o = rest[member];
    const int group = tileGroups[tile];
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const RigidGroup& r = groups[group];
    const int4 tiles = r.tiles;
    // a tile no block holds
    if (tile < tiles.x || tile >= tiles.x + tiles.y) return;
    float4 va = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vb = va;
    if (b >= 0) {
        const float4 state = body[b];
        float px = state.x * a.width - r.frame.x;
        float py = state.y * a.height - r.frame.y;
        if (!a.bounded) {
            px -= floorf(px / a.width + 0.5f) * a.width;
            py -= floorf(py / a.height + 0.5f) * a.height;
        }
        const float vx = state.z * a.width, vy = state.w * a.height;
        const float rx = r.motion.z, ry = r.motion.w;
        const float armx = rx * o.x - ry * o.y, army = ry * o.x + rx * o.y;
        va = make_float4(px, py, vx, vy);
        vb = make_float4(o.x * py - o.y * px, o.x * px + o.y * py, armx * vy - army * vx, 0.0f);
    }
    sa[lane] = va;
    sb[lane] = vb;
    tree_sum(sa, sb, lane);
    if (lane == 0) {
        partials[2 * tile] = sa[0];
        partials[2 * tile + 1] = sb[0];
    }
}

// S/rigid_reduce.comp:42-58: stopAgainst.
__device__ __forceinline__ void stop_against(float& vx, float& vy, float pushx, float pushy,
                                             float friction, float substepDt) {
    const float distance = sqrtf(pushx * pushx + pushy * pushy);
    if (distance == 0.0f) return;
    const float dx = pushx / distance, dy = pushy / distance;
    const float against = vx * dx + vy * dy;
    if (against >= 0.0f) return;
    const float stop = fminf(-against, distance / substepDt);
    vx = vx + dx * stop;
    vy = vy + dy * stop;
    const float along = vx * dx + vy * dy;
    const float slx = vx - along * dx, sly = vy - along * dy;
    const float slidingSpeed = sqrtf(slx * slx + sly * sly);
    if (slidingSpeed > 0.0f) {
        const float f = fminf(slidingSpeed, friction * stop) / slidingSpeed;
        vx = vx - slx * f;
        vy = vy - sly * f;
    }
}

// S/rigid_reduce.comp:78-100: settleContacts. The pairs a block owns are its
// partners from itself on, the four walls and the loose bodies.
__device__ void settle_contacts(int* __restrict__ memory, int group, int lane, int groupCount) {
    int* impacts = memory;
    int* nextImpacts = memory + RIGID_CONTACT_KEYS;
    int* contactStates = memory + 2 * RIGID_CONTACT_KEYS;
    int* idlePasses = memory + 3 * RIGID_CONTACT_KEYS;
    int* pairContacts = memory + 4 * RIGID_CONTACT_KEYS;
    const int partners = groupCount - group;
    for (int index = lane; index < partners + 5; index += RIGID_TILE) {
        const int partner = index < partners ? group + index : RIGID_MAXIMUM_GROUPS + index - partners;
        const int key = group * RIGID_CONTACT_PARTNERS + partner;
        const int state = contactStates[key];
        if (state == CONTACT_PUSHING) {
            const int speed = max(impacts[key] & IMPACT_SPEED_MASK, nextImpacts[key]);
            const int age = (impacts[key] >> IMPACT_AGE_SHIFT) + 1;
            impacts[key] = speed == 0 || age > IMPACT_PUSHING_PASSES ? 0 : speed | (age << IMPACT_AGE_SHIFT);
            idlePasses[key] = 0;
        } else if (state == CONTACT_TOUCHING) {
            impacts[key] = 0;
            idlePasses[key] = 0;
        } else if (impacts[key] != 0) {
            const int idle = idlePasses[key] + 1;
            idlePasses[key] = idle;
            if (idle >= IMPACT_IDLE_PASSES) {
                impacts[key] = 0;
                idlePasses[key] = 0;
            }
        }
        nextImpacts[key] = 0;
        contactStates[key] = 0;
        pairContacts[key] = 0;
    }
}

// S/rigid_reduce.comp:102-201: fit. One workgroup per block: the tile sums
// added in tile order (lane i takes tiles i, i + 256, ...) and over the
// lanes by the tree, then the fit on lane 0.
__global__ void __launch_bounds__(RIGID_TILE) k_rigid_fit(
    RigidGroup* __restrict__ groups, int* __restrict__ memory, const float* __restrict__ shapes,
    const float4* __restrict__ partials, RigidArgs a) {
    __shared__ float4 sa[RIGID_TILE], sb[RIGID_TILE];
    const int group = blockIdx.x, lane = threadIdx.x;
    const int4 tiles = groups[group].tiles;
    // an index no block was glued at
    if (tiles.y == 0) return;
    float4 va = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vb = va;
    for (int tile = tiles.x + lane; tile < tiles.x + tiles.y; tile += RIGID_TILE) {
        const float4 pa = partials[2 * tile], pb = partials[2 * tile + 1];
        va = make_float4(va.x + pa.x, va.y + pa.y, va.z + pa.z, va.w + pa.w);
        vb = make_float4(vb.x + pb.x, vb.y + pb.y, vb.z + pb.z, vb.w + pb.w);
    }
    sa[lane] = va;
    sb[lane] = vb;
    tree_sum(sa, sb, lane);
    if (a.contactPass) settle_contacts(memory, group, lane, a.groupCount);
    if (lane != 0) return;

    RigidGroup r = groups[group];
    const float mass = r.scalars.z;
    // broken up entirely
    if (mass <= 0.0f) return;
    const float4 A = sa[0], B = sb[0];
    float cx = A.x / mass, cy = A.y / mass;
    float vx = A.z / mass, vy = A.w / mass;
    // best-fit rotation of the rest shape onto the current positions
    const float fx = B.y, fy = B.x;
    const float fitLength = sqrtf(fx * fx + fy * fy);
    float rx = r.motion.z, ry = r.motion.w;
    if (fitLength > 0.0f) {
        rx = fx / fitLength;
        ry = fy / fitLength;
    }
    const float angularMomentum = B.z;
    const float inertia = r.scalars.y;
    const float angularVelocity = inertia > 0.0f ? angularMomentum / inertia : 0.0f;

    cx += r.frame.x;
    cy += r.frame.y;
    // out of the blocks it was pressed into, position only
    cx += (float)(r.guard.x - r.guard.z) / GUARD_SCALE;
    cy += (float)(r.guard.y - r.guard.w) / GUARD_SCALE;
    r.guard = make_int4(0, 0, 0, 0);
    // off bodies a wall holds
    const float sx = (float)(r.support.x - r.support.z) / GUARD_SCALE;
    const float sy = (float)(r.support.y - r.support.w) / GUARD_SCALE;
    r.support = make_int4(0, 0, 0, 0);
    cx += sx;
    cy += sy;
    if (a.contactPass) stop_against(vx, vy, sx, sy, 0.0f, a.substepDt);
    if (!a.bounded) {
        cx -= floorf(cx / a.width) * a.width;
        cy -= floorf(cy / a.height) * a.height;
    } else {
        // the hull reaches a wall first: back inside, stopped as the wall would
        float lowx = 1e30f, lowy = 1e30f, highx = -1e30f, highy = -1e30f;
        for (int p = 0; p < r.tiles.w; p++) {
            const float ox = shapes[r.tiles.z + 2 * p], oy = shapes[r.tiles.z + 2 * p + 1];
            const float armx = rx * ox - ry * oy, army = ry * ox + rx * oy;
            lowx = fminf(lowx, armx);
            lowy = fminf(lowy, army);
            highx = fmaxf(highx, armx);
            highy = fmaxf(highy, army);
        }
        float shx = fmaxf(a.bodyRadius - (cx + lowx), 0.0f) - fmaxf(cx + highx + a.bodyRadius - a.width, 0.0f);
        float shy = fmaxf(a.bodyRadius - (cy + lowy), 0.0f) - fmaxf(cy + highy + a.bodyRadius - a.height, 0.0f);
        // a block wider than the screen stays where it is
        shx *= highx - lowx + 2.0f * a.bodyRadius < a.width ? 1.0f : 0.0f;
        shy *= highy - lowy + 2.0f * a.bodyRadius < a.height ? 1.0f : 0.0f;
        cx += shx;
        cy += shy;
        if (a.contactPass) stop_against(vx, vy, shx, shy, a.blockFriction, a.substepDt);
    }
    r.frame = make_float4(cx, cy, cx, cy);
    r.motion = make_float4(vx, vy, rx, ry);
    r.scalars.x = angularVelocity;
    groups[group] = r;
}

// S/rigid_apply.comp:48-69: every member onto its block's rigid motion,
// x = c + R r0 and v = V + w x (R r0).
__global__ void __launch_bounds__(RIGID_TILE) k_rigid_apply(
    float4* __restrict__ body, const int* __restrict__ members, const float2* __restrict__ rest,
    const int* __restrict__ tileGroups, const RigidGroup* __restrict__ groups, RigidArgs a) {
    const int tile = blockIdx.x;
    const int member = tile * RIGID_TILE + (int)threadIdx.x;
    const int b = members[member];
    const float2 o = rest[member];
    const int group = tileGroups[tile];
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (b < 0) return;
    const RigidGroup& r = groups[group];
    const float rx = r.motion.z, ry = r.motion.w;
    const float armx = rx * o.x - ry * o.y, army = ry * o.x + rx * o.y;
    float px = (r.frame.z + armx) / a.width, py = (r.frame.w + army) / a.height;
    const float w = r.scalars.x;
    const float vx = r.motion.x + w * -army, vy = r.motion.y + w * armx;
    if (a.bounded) {
        px = fminf(fmaxf(px, 0.0f), MAXIMUM_UNIT);
        py = fminf(fmaxf(py, 0.0f), MAXIMUM_UNIT);
    } else {
        px = fractf(px);
        py = fractf(py);
    }
    body[b] = make_float4(px, py, vx / a.width, vy / a.height);
}

// S/rigid_adopt.comp:66-75: block index `group` starts afresh, its contact
// memory with every partner forgotten.
__global__ void __launch_bounds__(RIGID_TILE) k_rigid_adopt(int* __restrict__ memory, int group) {
    const int partner = blockIdx.x * RIGID_TILE + (int)threadIdx.x;
    if (partner >= RIGID_CONTACT_PARTNERS) return;
    const int key = min(group, partner) * RIGID_CONTACT_PARTNERS + max(group, partner);
    for (int plane = 0; plane < 5; plane++) memory[plane * RIGID_CONTACT_KEYS + key] = 0;
}

// S/rigid_remap.comp: members follow the bodies spatial sorting renamed. The
// sort moves the `active` bodies only: a member beyond them keeps its index.
__global__ void __launch_bounds__(RIGID_TILE) k_rigid_remap(int* __restrict__ members,
                                                           const uint32_t* __restrict__ dest, int count,
                                                           int active) {
    const int member = blockIdx.x * RIGID_TILE + (int)threadIdx.x;
    if (member >= count) return;
    const int b = members[member];
    if (b >= 0 && b < active) members[member] = (int)dest[b];
}

// Kotlin's total order on floats (Float.compareTo: -0 below +0) and its
// equality of data classes with float fields (bitwise).
bool float_before(float a, float b) {
    if (a < b) return true;
    return a == b && std::signbit(a) && !std::signbit(b);
}

bool float_same(float a, float b) {
    return a == b && std::signbit(a) == std::signbit(b);
}

// GlueShape.hullIndices (K/GlueShape.kt:158-190): the corners of the convex
// hull of the rest centres, counter-clockwise from the lowest x, found among
// the ends of each row of centres at one height.
std::vector<int> hull_indices(const std::vector<float>& xs, const std::vector<float>& ys) {
    std::vector<int> chain;
    if (xs.empty()) return chain;
    const float minY = *std::min_element(ys.begin(), ys.end());
    const float maxY = *std::max_element(ys.begin(), ys.end());
    // a row table over the block's height (the reference's IntArray per row)
    if (!((maxY - minY) / HULL_ROW_PIXELS < (float)(1 << 24)))
        throw ArgumentError("a block's members must lie within 167,772 px of one another");
    const int rows = (int)((maxY - minY) / HULL_ROW_PIXELS) + 1;
    std::vector<int> lowest(rows, -1), highest(rows, -1);
    for (int i = 0; i < (int)xs.size(); i++) {
        const int row = (int)((ys[i] - minY) / HULL_ROW_PIXELS);
        if (lowest[row] < 0 || xs[i] < xs[lowest[row]]) lowest[row] = i;
        if (highest[row] < 0 || xs[i] > xs[highest[row]]) highest[row] = i;
    }
    // both ends of every row, first of each distinct position kept
    std::vector<int> points;
    for (int row = 0; row < rows; row++) {
        if (lowest[row] < 0) continue;
        for (int idx : {lowest[row], highest[row]}) {
            bool seen = false;
            for (int p : points) seen = seen || (float_same(xs[p], xs[idx]) && float_same(ys[p], ys[idx]));
            if (!seen) points.push_back(idx);
        }
    }
    std::stable_sort(points.begin(), points.end(), [&](int p, int q) {
        if (float_before(xs[p], xs[q])) return true;
        return float_same(xs[p], xs[q]) && float_before(ys[p], ys[q]);
    });
    if (points.size() < 3) return points;
    auto turn = [&](int o, int p, int q) {
        return (double)(xs[p] - xs[o]) * (double)(ys[q] - ys[o]) -
               (double)(ys[p] - ys[o]) * (double)(xs[q] - xs[o]);
    };
    for (int pass = 0; pass < 2; pass++) {
        const size_t start = chain.size();
        for (size_t k = 0; k < points.size(); k++) {
            const int point = pass == 0 ? points[k] : points[points.size() - 1 - k];
            while (chain.size() >= start + 2 && turn(chain[chain.size() - 2], chain.back(), point) <= 0.0)
                chain.pop_back();
            chain.push_back(point);
        }
        chain.pop_back();
    }
    return chain;
}

}  // namespace

// RectangularBlock.offsets (K/RectangularBlock.kt:10-77): a rectangle filled
// with hexagonally packed bodies about its centre, row by row, each row from
// left to right, every body inside.
void block_offsets(float diameter, float width, float height, std::vector<float>& xy) {
    xy.clear();
    if (!(diameter > 0.0f)) throw ArgumentError("contact diameter must be positive");
    constexpr float EPSILON = 1e-4f;
    const float halfX = 0.5f * (width - diameter), halfY = 0.5f * (height - diameter);
    if (halfX < 0.0f || halfY < 0.0f) return;
    const float rowStep = diameter * (float)(std::sqrt(3.0) * 0.5);
    const int rows = (int)std::floor(2.0f * halfY / rowStep + EPSILON) + 1;
    const int evenColumns = (int)std::floor(2.0f * halfX / diameter + EPSILON) + 1;
    const float oddSpan = 2.0f * halfX - 0.5f * diameter;
    const int oddColumns = oddSpan < -EPSILON * diameter ? 0 : (int)std::floor(oddSpan / diameter + EPSILON) + 1;
    for (int row = 0; row < rows; row++) {
        const bool odd = (row & 1) == 1;
        const int columns = odd ? oddColumns : evenColumns;
        const float startX = -halfX + (odd ? 0.5f * diameter : 0.0f);
        const float y = -halfY + (float)row * rowStep;
        for (int column = 0; column < columns; column++) {
            xy.push_back(startX + (float)column * diameter);
            xy.push_back(y);
        }
    }
}

// ---------------------------------------------------------------------------
// Engine: the blocks of simulation 0

void Engine::rigid_allocate() {
    if (rigid_.groups) return;
    const HostSim& h = sims_[0];
    // K/RigidBlocks.kt:90-91: every block may leave up to a tile's worth of padding
    const long long cap = h.cap;
    rigid_.memberCap = (int)((cap + (long long)RIGID_MAXIMUM_GROUPS * (RIGID_TILE - 1) + RIGID_TILE - 1) /
                             RIGID_TILE * RIGID_TILE);
    rigid_.tileCap = rigid_.memberCap / RIGID_TILE;
    // the hulls: at most one corner per member, two values each
    rigid_.valueCap = std::max<long long>(1LL << 18, 2 * cap);
    auto alloc = [&](void** p, size_t bytes) {
        check(hipMalloc(p, bytes), "rigid allocation");
        check(hipMemset(*p, 0, bytes), "rigid clear");
    };
    alloc((void**)&rigid_.groups, sizeof(RigidGroup) * RIGID_MAXIMUM_GROUPS);
    alloc((void**)&rigid_.memory, sizeof(int) * 5 * (size_t)RIGID_CONTACT_KEYS);
    alloc((void**)&rigid_.shapes, sizeof(float) * (size_t)rigid_.valueCap);
    alloc((void**)&rigid_.members, sizeof(int) * (size_t)rigid_.memberCap);
    alloc((void**)&rigid_.rest, sizeof(float2) * (size_t)rigid_.memberCap);
    alloc((void**)&rigid_.tileGroups, sizeof(int) * (size_t)rigid_.tileCap);
    alloc((void**)&rigid_.partials, sizeof(float4) * 2 * (size_t)rigid_.tileCap);
    // no block lists a slot yet
    std::vector<int> none(rigid_.memberCap, -1);
    check(hipMemcpy(rigid_.members, none.data(), sizeof(int) * none.size(), hipMemcpyHostToDevice),
          "rigid members");
}

void Engine::rigid_free() {
    void* ptrs[] = {rigid_.groups, rigid_.memory, rigid_.shapes, rigid_.members,
                    rigid_.rest, rigid_.tileGroups, rigid_.partials};
    for (void* p : ptrs)
        if (p) hipFree(p);
    rigid_ = RigidState{};
}

int Engine::rigid_glue(int firstBody, int count, const float* offsets, float centerX, float centerY) {
    if (sims_.size() != 1) throw ArgumentError("glued blocks need a pipeline of one simulation");
    const HostSim& h = sims_[0];
    if (count <= 0 || firstBody < 0 || (long long)firstBody + count > h.cap)
        throw ArgumentError("block bodies must lie within capacity");
    rigid_allocate();
    // the uploads below are blocking copies beside the handle's stream: let its work end first
    synchronize();
    const int group = rigid_.groupCount;   // nextGroup: no block is ever freed here
    const int tiles = (count + RIGID_TILE - 1) / RIGID_TILE;
    // canAdd (K/RigidBlocks.kt:276-278)
    if (group >= RIGID_MAXIMUM_GROUPS) throw ArgumentError("no group index left for a block");
    if (rigid_.tileCount + tiles > rigid_.tileCap) throw ArgumentError("no member slots left for a block");

    // prepare (K/RigidBlocks.kt:751-777): offsets about their centroid, summed in double
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < count; i++) {
        sx += (double)offsets[2 * i];
        sy += (double)offsets[2 * i + 1];
    }
    const float mx = (float)(sx / count), my = (float)(sy / count);
    std::vector<float> restX(count), restY(count);
    double inertia = 0.0;
    float minX = offsets[0], maxX = offsets[0], minY = offsets[1], maxY = offsets[1];
    for (int i = 0; i < count; i++) {
        const float ox = offsets[2 * i], oy = offsets[2 * i + 1];
        const float x = ox - mx, y = oy - my;
        restX[i] = x;
        restY[i] = y;
        const float term = x * x + y * y;
        inertia += (double)term;
        minX = std::min(minX, ox);
        maxX = std::max(maxX, ox);
        minY = std::min(minY, oy);
        maxY = std::max(maxY, oy);
    }
    const std::vector<int> corners = hull_indices(restX, restY);
    const long long values = 2LL * (long long)corners.size();
    if (rigid_.valueTop + values > rigid_.valueCap) throw ArgumentError("no shape values left for a block");

    // upload (K/RigidBlocks.kt:779-857), with no distance field: the block's
    // region of the shape values holds only its hull
    const int firstTile = rigid_.tileCount;
    const int entries = tiles * RIGID_TILE;
    std::vector<int> memberData(entries, -1);
    std::vector<float2> restData(entries, make_float2(0.0f, 0.0f));
    for (int i = 0; i < count; i++) {
        memberData[i] = firstBody + i;
        restData[i] = make_float2(restX[i], restY[i]);
    }
    std::vector<float> hull(values);
    for (size_t k = 0; k < corners.size(); k++) {
        hull[2 * k] = restX[corners[k]];
        hull[2 * k + 1] = restY[corners[k]];
    }
    std::vector<int> tileData(tiles, group);
    RigidGroup g{};
    // the first projection fits the bodies' velocity
    g.frame = make_float4(centerX, centerY, centerX, centerY);
    g.motion = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    g.scalars = make_float4(0.0f, (float)inertia, (float)count, 0.0f);
    g.tiles = make_int4(firstTile, tiles, (int)rigid_.valueTop, (int)corners.size());
    g.shape = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    g.grid = make_int4((int)rigid_.valueTop, 0, 0, 0);
    g.bounds = make_float4(minX - mx, maxX - mx, minY - my, maxY - my);
    g.guard = make_int4(0, 0, 0, 0);
    g.support = make_int4(0, 0, 0, 0);
    g.pose = make_float4(centerX, centerY, 1.0f, 0.0f);

    check(hipMemcpy(rigid_.members + (size_t)firstTile * RIGID_TILE, memberData.data(),
                    sizeof(int) * entries, hipMemcpyHostToDevice), "rigid members");
    check(hipMemcpy(rigid_.rest + (size_t)firstTile * RIGID_TILE, restData.data(),
                    sizeof(float2) * entries, hipMemcpyHostToDevice), "rigid rest");
    check(hipMemcpy(rigid_.tileGroups + firstTile, tileData.data(), sizeof(int) * tiles,
                    hipMemcpyHostToDevice), "rigid tiles");
    if (values > 0)
        check(hipMemcpy(rigid_.shapes + rigid_.valueTop, hull.data(), sizeof(float) * values,
                        hipMemcpyHostToDevice), "rigid hull");
    check(hipMemcpy(rigid_.groups + group, &g, sizeof(RigidGroup), hipMemcpyHostToDevice), "rigid group");
    // adoptGroup (K/RigidBlocks.kt:859-870): the contact memory of the index
    // with every partner starts empty
    launch(k_rigid_adopt, dim3((RIGID_CONTACT_PARTNERS + RIGID_TILE - 1) / RIGID_TILE), dim3(RIGID_TILE), 0,
           st_, rigid_.memory, group);
    check(hipStreamSynchronize(st_), "rigid adopt");
    rigid_.groupCount = group + 1;
    rigid_.tileCount = firstTile + tiles;
    rigid_.valueTop += values;
    return group;
}

int Engine::spawn_block(int firstBody, float minX, float minY, float maxX, float maxY, float density,
                        float velocityX, float velocityY, float spin) {
    // K/GravityPipeline.kt:647-672 (no mask), glueBlock :692-717, uploadBodies :733-784
    if (sims_.size() != 1) throw ArgumentError("glued blocks need a pipeline of one simulation");
    const HostSim& h = sims_[0];
    if (!(density >= 0.05f && density <= 10.0f)) throw ArgumentError("block density must be in 0.05..10.0");
    const float diameter = h.cfg.contact_diameter_pixels;
    std::vector<float> lattice;
    block_offsets(diameter, maxX - minX, maxY - minY, lattice);
    const int count = (int)(lattice.size() / 2);
    if (count == 0 || !(firstBody >= 0 && firstBody <= h.cap && count <= h.cap - firstBody)) return 0;
    const float centerX = 0.5f * (minX + maxX), centerY = 0.5f * (minY + maxY);
    // the rigid fit needs offsets about the centre of mass
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < count; i++) {
        sx += (double)lattice[2 * i];
        sy += (double)lattice[2 * i + 1];
    }
    const float meanX = (float)sx / (float)count, meanY = (float)sy / (float)count;
    std::vector<float> offsets(lattice.size());
    for (int i = 0; i < count; i++) {
        offsets[2 * i] = lattice[2 * i] - meanX;
        offsets[2 * i + 1] = lattice[2 * i + 1] - meanY;
    }
    // canAdd (K/RigidBlocks.kt:276-278), the hull standing for the shape's values
    rigid_allocate();
    const int tiles = (count + RIGID_TILE - 1) / RIGID_TILE;
    if (rigid_.groupCount >= RIGID_MAXIMUM_GROUPS || rigid_.tileCount + tiles > rigid_.tileCap ||
        rigid_.valueTop + 2LL * count > rigid_.valueCap)
        return 0;
    const int group = rigid_.groupCount;
    const float cx = centerX + meanX, cy = centerY + meanY;
    const float W = (float)h.cfg.viewport_width, H = (float)h.cfg.viewport_height;
    auto place = [&](float v, float extent, int size) {
        if (h.bounded) return std::min(std::max(v, 0.0f), extent * MAXIMUM_UNIT);
        return v - std::floor(v / (float)size) * (float)size;
    };
    std::vector<float> bodies((size_t)4 * count), mats((size_t)4 * count);
    for (int i = 0; i < count; i++) {
        const float ox = offsets[2 * i], oy = offsets[2 * i + 1];
        bodies[4 * i] = place(cx + ox, W, h.cfg.viewport_width) / W;
        bodies[4 * i + 1] = place(cy + oy, H, h.cfg.viewport_height) / H;
        bodies[4 * i + 2] = (velocityX - spin * oy) / W;
        bodies[4 * i + 3] = (velocityY + spin * ox) / H;
        mats[4 * i] = density;
        mats[4 * i + 1] = h.cfg.block_friction;
        mats[4 * i + 2] = 1.0f;
        mats[4 * i + 3] = (float)group;
    }
    upload(0, firstBody, count, bodies.data(), mats.data());
    rigid_glue(firstBody, count, offsets.data(), cx, cy);
    return count;
}

void Engine::rigid_project(bool contacts) {
    // K/RigidBlocks.kt:309-357 for blocks that nothing breaks
    if (rigid_.groupCount == 0) return;
    normalize();
    const HostSim& h = sims_[0];
    RigidArgs a{(float)h.cfg.viewport_width, (float)h.cfg.viewport_height, h.bounded, contacts ? 1 : 0,
                rigid_.groupCount, h.desc.subDt, 0.5f * h.cfg.contact_diameter_pixels, h.cfg.block_friction};
    float4* body = body_[bodyPar_];
    launch(k_rigid_reduce, dim3(rigid_.tileCount), dim3(RIGID_TILE), 0, st_, body, rigid_.members,
           rigid_.rest, rigid_.tileGroups, rigid_.groups, rigid_.partials, a);
    launch(k_rigid_fit, dim3(rigid_.groupCount), dim3(RIGID_TILE), 0, st_, rigid_.groups, rigid_.memory,
           rigid_.shapes, rigid_.partials, a);
    launch(k_rigid_apply, dim3(rigid_.tileCount), dim3(RIGID_TILE), 0, st_, body, rigid_.members,
           rigid_.rest, rigid_.tileGroups, rigid_.groups, a);
    check(hipGetLastError(), "rigid projection");
}

void Engine::rigid_remap() {
    // K/RigidBlocks.kt:391-399, after every spatial sort
    if (rigid_.groupCount == 0) return;
    const int count = rigid_.tileCount * RIGID_TILE;
    launch(k_rigid_remap, dim3(rigid_.tileCount), dim3(RIGID_TILE), 0, st_, rigid_.members,
           (const uint32_t*)dest_, count, sims_[0].n);
}

void Engine::rigid_size(int* out) const {
    out[0] = rigid_.groupCount;
    out[1] = rigid_.tileCount;
    out[2] = rigid_.memberCap;
}

void Engine::rigid_access(int what, long long first, long long count, void* host, bool write) {
    if (count < 0 || first < 0) throw ArgumentError("range out of bounds");
    if (count == 0) return;
    rigid_allocate();
    char* base = nullptr;
    long long limit = 0;
    size_t unit = 4;
    switch (what) {
        case GS_RIGID_GROUPS: base = (char*)rigid_.groups; limit = RIGID_MAXIMUM_GROUPS * 40LL; break;
        case GS_RIGID_CONTACT_MEMORY: base = (char*)rigid_.memory; limit = 5LL * RIGID_CONTACT_KEYS; break;
        case GS_RIGID_MEMBERS: base = (char*)rigid_.members; limit = rigid_.memberCap; break;
        case GS_RIGID_REST: base = (char*)rigid_.rest; limit = 2LL * rigid_.memberCap; break;
        case GS_RIGID_SHAPES: base = (char*)rigid_.shapes; limit = rigid_.valueCap; break;
        case GS_RIGID_TILE_GROUPS: base = (char*)rigid_.tileGroups; limit = rigid_.tileCap; break;
        default: throw ArgumentError("unknown rigid buffer");
    }
    if (write && what != GS_RIGID_GROUPS && what != GS_RIGID_CONTACT_MEMORY)
        throw ArgumentError("only the groups and the contact memory are writable");
    if (first + count > limit) throw ArgumentError("range out of bounds");
    // the tile and hull ranges of a block index the other buffers: the library's own
    if (write && what == GS_RIGID_GROUPS)
        for (long long w = first; w < first + count; w++)
            if (w % 40 >= 12 && w % 40 < 16) throw ArgumentError("a block's tile and hull words are not writable");
    synchronize();
    if (write)
        check(hipMemcpy(base + first * unit, host, count * unit, hipMemcpyHostToDevice), "rigid write");
    else
        check(hipMemcpy(host, base + first * unit, count * unit, hipMemcpyDeviceToHost), "rigid read");
}

}  // namespace gs
