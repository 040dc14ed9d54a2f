// wire.cpp -- parry's Bvh and TriMesh as serialized bytes (host only, no GPU).
//
// parry derives serde's Serialize/Deserialize for Bvh (bvh_tree.rs:1748-1762:
// nodes, parents, leaf_node_indices), BvhNodeWide/BvhNode/BvhNodeData/
// BvhNodeIndex (bvh_tree.rs:148-156, 252-266, 447-467, 1243-1250), VecMap
// (utils/vec_map.rs:73-88: n, v) and TriMesh (trimesh.rs:510-528: bvh,
// vertices, indices, pseudo_normals, topology, connected_components, flags),
// so the bytes are those fields in declaration order as the chosen serializer
// writes them.  The format here is bincode 1.x with its default options
// (`bincode::serialize`): little-endian, fixed-width integers, usize as u64,
// a Vec as a u64 length then its elements, an Option as a u8 tag (0 None,
// 1 Some) then the value, structs, newtypes, tuples and fixed arrays as their
// fields with nothing in between.  nalgebra's Point3<f32> (its coordinates,
// an ArrayStorage serialized as a 3-tuple) and [u32; 3] are 12 bytes, a
// BvhNode 32 (mins, children, maxs, data), a BvhNodeWide 64 -- the in-memory
// records -- a BvhNodeIndex a u64 ((wide << 1) | is_right), TriMeshFlags its u16.
//
// A Bvh built by Bvh::from_leaves / TriMesh::new has, besides its depth-first
// nodes: parents[w] = the slot (wide, side) holding wide node w, parents[0] =
// 0 (bvh_refit.rs:259-320, bvh_tree.rs:1914-1946); leaf_node_indices with
// n = #leaves and v[leaf] = Some(slot of the leaf) for leaf ids 0..n-1.
// Encoding derives both from the nodes; decoding checks them against the
// nodes.  TriMesh::new leaves pseudo_normals, topology and connected_components
// None and flags empty (trimesh.rs:607-612, 718-746); meshes with any of them
// set are rejected by the decoder (PQ_ERR_UNSUPPORTED).
#include <cstring>
#include <string>
#include <vector>

#include "pq_internal.h"

using namespace pq;

namespace {

constexpr uint32_t kLeafCountMask = 0x3fffffffu;

struct Writer {
    uint8_t *out;
    size_t cap, len = 0;
    void raw(const void *p, size_t n) {
        if (out && len + n <= cap) std::memcpy(out + len, p, n);
        len += n;
    }
    void u8(uint8_t v) { raw(&v, 1); }
    void u16(uint16_t v) { raw(&v, 2); }
    void u64(uint64_t v) { raw(&v, 8); }
};

struct Reader {
    const uint8_t *p;
    size_t size, pos = 0;
    bool ok = true;
    bool raw(void *dst, size_t n) {
        if (!ok || n > size - pos) return ok = false;
        std::memcpy(dst, p + pos, n);
        pos += n;
        return true;
    }
    uint8_t u8() {
        uint8_t v = 0;
        raw(&v, 1);
        return v;
    }
    uint16_t u16() {
        uint16_t v = 0;
        raw(&v, 2);
        return v;
    }
    uint64_t u64() {
        uint64_t v = 0;
        raw(&v, 8);
        return v;
    }
    // A length prefix whose elements of `elem` bytes must fit in what is left.
    uint64_t len(size_t elem) {
        const uint64_t n = u64();
        if (ok && elem && n > (size - pos) / elem) ok = false;
        return ok ? n : 0;
    }
};

// The parents and leaf slots of a depth-first tree, as parry's Bvh keeps them.
// Returns false if the nodes do not form such a tree over leaves 0..n-1.
bool tree_slots(const pq_bvh_node_wide *nodes, uint32_t nn, std::vector<uint64_t> &parents,
                std::vector<uint64_t> &leaf_slot, std::string &why) {
    parents.assign(nn, 0);
    std::vector<uint32_t> leaf_count(nn, 0);
    std::vector<uint8_t> seen_node(nn, 0);
    uint64_t leaves = 0;
    std::vector<std::pair<uint32_t, uint64_t>> leaf_at;  // (leaf id, slot)
    for (uint32_t w = 0; w < nn; ++w) {
        const pq_bvh_node *side[2] = {&nodes[w].left, &nodes[w].right};
        for (uint32_t s = 0; s < 2; ++s) {
            const uint32_t lc = side[s]->data & kLeafCountMask;
            const uint64_t slot = ((uint64_t)w << 1) | s;
            if (lc == 1) {
                leaf_at.push_back({side[s]->children, slot});
                ++leaves;
            } else if (lc > 1) {
                const uint32_t c = side[s]->children;
                if (c <= w || c >= nn || seen_node[c]) {
                    why = "wide node " + std::to_string(w) + " has an invalid child index";
                    return false;
                }
                seen_node[c] = 1;
                parents[c] = slot;
            } else if (!(w == 0 && s == 1 && nn == 1)) {
                why = "empty child slot outside a one-leaf root";
                return false;
            }
        }
    }
    for (uint32_t w = 1; w < nn; ++w)
        if (!seen_node[w]) {
            why = "wide node " + std::to_string(w) + " has no parent";
            return false;
        }
    leaf_slot.assign(leaves, UINT64_MAX);
    for (const auto &ls : leaf_at) {
        if (ls.first >= leaves || leaf_slot[ls.first] != UINT64_MAX) {
            why = "leaf ids are not 0..n-1, each once";
            return false;
        }
        leaf_slot[ls.first] = ls.second;
    }
    return true;
}

void write_nodes(Writer &w, const pq_bvh_node_wide *nodes, uint32_t nn) {
    w.u64(nn);
    w.raw(nodes, (size_t)nn * sizeof(pq_bvh_node_wide));
}

int encode_bvh(Writer &w, const pq_bvh_node_wide *nodes, uint32_t nn, const char *who) {
    std::vector<uint64_t> parents, leaf_slot;
    std::string why;
    if (nn && !tree_slots(nodes, nn, parents, leaf_slot, why)) return fail(PQ_ERR_INVALID, std::string(who) + ": " + why);
    write_nodes(w, nodes, nn);
    w.u64(parents.size());
    for (uint64_t p : parents) w.u64(p);
    w.u64(leaf_slot.size());  // VecMap::n
    w.u64(leaf_slot.size());  // VecMap::v
    for (uint64_t s : leaf_slot) {
        w.u8(1);
        w.u64(s);
    }
    return PQ_OK;
}

int decode_bvh(Reader &r, std::vector<pq_bvh_node_wide> &nodes, const char *who) {
    const uint64_t nn = r.len(sizeof(pq_bvh_node_wide));
    if (!r.ok || nn > UINT32_MAX) return fail(PQ_ERR_INVALID, std::string(who) + ": truncated or oversized node list");
    nodes.resize(nn);
    r.raw(nodes.data(), nn * sizeof(pq_bvh_node_wide));
    const uint64_t np = r.len(8);
    std::vector<uint64_t> parents(np);
    for (auto &p : parents) p = r.u64();
    const uint64_t n_some = r.u64();
    const uint64_t nv = r.len(1);
    std::vector<uint64_t> slots(nv, UINT64_MAX);
    uint64_t some = 0;
    for (auto &s : slots) {
        const uint8_t tag = r.u8();
        if (tag == 1) {
            s = r.u64();
            ++some;
        } else if (tag != 0) {
            r.ok = false;
        }
    }
    if (!r.ok) return fail(PQ_ERR_INVALID, std::string(who) + ": truncated or malformed Bvh");
    std::vector<uint64_t> want_parents, want_slots;
    std::string why;
    if (nn && !tree_slots(nodes.data(), (uint32_t)nn, want_parents, want_slots, why))
        return fail(PQ_ERR_INVALID, std::string(who) + ": " + why);
    if (parents != want_parents || some != n_some || slots != want_slots)
        return fail(PQ_ERR_INVALID, std::string(who) + ": parents or leaf_node_indices disagree with the nodes");
    return PQ_OK;
}

}  // namespace

extern "C" {

int pq_bvh_encode(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint8_t *out, size_t capacity, size_t *size) {
    if (!size || (num_nodes && !nodes)) return fail(PQ_ERR_INVALID, "pq_bvh_encode: null argument");
    Writer w{out, out ? capacity : 0};
    const int rc = encode_bvh(w, nodes, num_nodes, "pq_bvh_encode");
    if (rc != PQ_OK) return rc;
    *size = w.len;
    if (w.len > capacity || !out) return fail(PQ_ERR_CAPACITY, "pq_bvh_encode: needs " + std::to_string(w.len) + " bytes");
    return PQ_OK;
}

int pq_bvh_decode(const uint8_t *bytes, size_t size, pq_bvh_node_wide *nodes, uint32_t capacity, uint32_t *num_nodes) {
    if (!num_nodes || (size && !bytes)) return fail(PQ_ERR_INVALID, "pq_bvh_decode: null argument");
    Reader r{bytes, size};
    std::vector<pq_bvh_node_wide> v;
    const int rc = decode_bvh(r, v, "pq_bvh_decode");
    if (rc != PQ_OK) return rc;
    if (r.pos != size) return fail(PQ_ERR_INVALID, "pq_bvh_decode: trailing bytes");
    *num_nodes = (uint32_t)v.size();
    if (v.size() > capacity || (!nodes && !v.empty()))
        return fail(PQ_ERR_CAPACITY, "pq_bvh_decode: needs " + std::to_string(v.size()) + " nodes");
    if (!v.empty()) std::memcpy(nodes, v.data(), v.size() * sizeof(pq_bvh_node_wide));
    return PQ_OK;
}

int pq_trimesh_encode(const float *vertices, uint32_t num_vertices, const uint32_t *indices, uint32_t num_triangles,
                      const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint8_t *out, size_t capacity,
                      size_t *size) {
    if (!size || (num_vertices && !vertices) || (num_triangles && !indices) || (num_nodes && !nodes))
        return fail(PQ_ERR_INVALID, "pq_trimesh_encode: null argument");
    Writer w{out, out ? capacity : 0};
    const int rc = encode_bvh(w, nodes, num_nodes, "pq_trimesh_encode");
    if (rc != PQ_OK) return rc;
    w.u64(num_vertices);
    w.raw(vertices, (size_t)num_vertices * 12);
    w.u64(num_triangles);
    w.raw(indices, (size_t)num_triangles * 12);
    w.u8(0);   // pseudo_normals: None
    w.u8(0);   // topology: None
    w.u8(0);   // connected_components: None
    w.u16(0);  // flags: empty
    *size = w.len;
    if (w.len > capacity || !out)
        return fail(PQ_ERR_CAPACITY, "pq_trimesh_encode: needs " + std::to_string(w.len) + " bytes");
    return PQ_OK;
}

int pq_trimesh_decode(const uint8_t *bytes, size_t size, float *vertices, uint32_t vertex_capacity,
                      uint32_t *num_vertices, uint32_t *indices, uint32_t triangle_capacity, uint32_t *num_triangles,
                      pq_bvh_node_wide *nodes, uint32_t node_capacity, uint32_t *num_nodes) {
    if (!num_vertices || !num_triangles || !num_nodes || (size && !bytes))
        return fail(PQ_ERR_INVALID, "pq_trimesh_decode: null argument");
    Reader r{bytes, size};
    std::vector<pq_bvh_node_wide> v;
    int rc = decode_bvh(r, v, "pq_trimesh_decode");
    if (rc != PQ_OK) return rc;
    const uint64_t nv = r.len(12);
    std::vector<float> vert(nv * 3);
    r.raw(vert.data(), nv * 12);
    const uint64_t nt = r.len(12);
    std::vector<uint32_t> idx(nt * 3);
    r.raw(idx.data(), nt * 12);
    const uint8_t pn = r.u8(), topo = r.u8(), cc = r.u8();
    const uint16_t flags = r.u16();
    if (!r.ok || nv > UINT32_MAX || nt > UINT32_MAX) return fail(PQ_ERR_INVALID, "pq_trimesh_decode: truncated TriMesh");
    if (pn > 1 || topo > 1 || cc > 1) return fail(PQ_ERR_INVALID, "pq_trimesh_decode: malformed Option tag");
    if (pn || topo || cc || flags)
        return fail(PQ_ERR_UNSUPPORTED, "pq_trimesh_decode: pseudo-normals, topology, connected components and "
                                        "flags are not part of this path (TriMesh::new leaves them unset)");
    if (r.pos != size) return fail(PQ_ERR_INVALID, "pq_trimesh_decode: trailing bytes");
    for (uint32_t k : idx)
        if (k >= nv) return fail(PQ_ERR_INVALID, "pq_trimesh_decode: vertex index out of range");
    std::vector<uint64_t> parents, slots;
    std::string why;
    if (!v.empty() && (!tree_slots(v.data(), (uint32_t)v.size(), parents, slots, why) || slots.size() != nt))
        return fail(PQ_ERR_INVALID, "pq_trimesh_decode: the Bvh's leaves are not the triangles 0..n-1");
    *num_vertices = (uint32_t)nv;
    *num_triangles = (uint32_t)nt;
    *num_nodes = (uint32_t)v.size();
    if (nv > vertex_capacity || nt > triangle_capacity || v.size() > node_capacity || (nv && !vertices) ||
        (nt && !indices) || (!v.empty() && !nodes))
        return fail(PQ_ERR_CAPACITY, "pq_trimesh_decode: needs " + std::to_string(nv) + " vertices, " +
                                         std::to_string(nt) + " triangles, " + std::to_string(v.size()) + " nodes");
    if (nv) std::memcpy(vertices, vert.data(), nv * 12);
    if (nt) std::memcpy(indices, idx.data(), nt * 12);
    if (!v.empty()) std::memcpy(nodes, v.data(), v.size() * sizeof(pq_bvh_node_wide));
    return PQ_OK;
}

}  // extern "C"
