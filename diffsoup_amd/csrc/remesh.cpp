// Longest-edge bisection of a triangle soup, on the host.
//
// Between training phases the soup is refined by repeatedly bisecting the
// longest edge in the whole soup (src/remesh.cpp:118-156) or the longest edge
// as seen through one camera (src/remesh_clip.cpp:200-244).  Bisecting
// triangle t = (a, b, c) across edge (a, b) rewrites t as (a, m, c') and
// appends (m', b, c): the soup stays a soup, so the split adds three vertices
// (two copies of the split point m and a copy of c) and one triangle.  Every
// triangle remembers the source triangle it descends from and whether it is
// still that triangle unchanged.
//
// The order of the output arrays is the order in which splits happen, which
// a max-heap keyed by squared edge length decides.  Entries carry the
// triangle's edit epoch at push time; an entry whose triangle has been edited
// since is stale and skipped.  Many keys tie exactly (the two halves of a
// split edge, the median shared by both children), so the heap must resolve
// ties as the reference's std::priority_queue does: the same container, the
// same comparison on the length alone and the same sequence of pushes and
// pops give the same order.
//
// Arithmetic follows the reference operation by operation (built with
// -ffp-contract=off): the outputs are bit-identical to it.
#include <cstdint>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "diffsoup_hip.h"

namespace ds {
int remesh_fail(int code, const char* msg);   // records the message for ds_last_error
}

struct ds_soup_split {
    std::vector<float> xyz;              // 3 per vertex
    std::vector<int> corner;             // 3 per triangle
    std::vector<int> source;             // source triangle of each triangle
    std::vector<unsigned char> pristine; // 1 while a triangle is its source unchanged
    std::vector<int> epoch;              // bumped whenever the triangle is rewritten
};

namespace ds {
namespace {

struct Candidate {
    float len2;
    int tri;
    int edge;   // 0: (v0, v1), 1: (v1, v2), 2: (v2, v0)
    int epoch;
};

// Max-heap on the length only: ties resolve by the heap's own sift order.
struct ShorterFirst {
    bool operator()(const Candidate& x, const Candidate& y) const { return x.len2 < y.len2; }
};
using EdgeQueue = std::priority_queue<Candidate, std::vector<Candidate>, ShorterFirst>;

// World-space metric (src/remesh.h:43-51, src/remesh.cpp:22-26): squared
// Euclidean length, split at the midpoint.  Every edge is queued.
struct WorldMetric {
    static constexpr bool kSkipZero = false;
    float len2(const float* A, const float* B) const {
        const float dx = B[0] - A[0], dy = B[1] - A[1], dz = B[2] - A[2];
        return dx * dx + dy * dy + dz * dz;
    }
    void split_point(const float* A, const float* B, float* M) const {
        for (int k = 0; k < 3; ++k) M[k] = 0.5f * (A[k] + B[k]);
    }
};

// Screen-space metric for one camera (src/remesh_clip.cpp:6-63, 92-126):
// lengths in NDC with x scaled by W/H (one unit = one image height); an edge
// with an end outside the NDC cube or behind the eye has length 0 and is never
// queued.  The split point is where the edge's projection crosses the NDC
// midpoint of its projected ends (perspective-correct in world space).
struct ScreenMetric {
    static constexpr bool kSkipZero = true;
    float m[16];
    float aspect;

    void clip(const float* P, float* c) const {
        for (int r = 0; r < 4; ++r) c[r] = m[4 * r] * P[0] + m[4 * r + 1] * P[1] + m[4 * r + 2] * P[2] + m[4 * r + 3];
    }
    // NDC of a clip position; false when w <= 0 (the NDC value is then 0)
    static bool ndc(const float* c, float* n) {
        const bool ok = c[3] > 0.0f;
        for (int k = 0; k < 3; ++k) n[k] = ok ? c[k] / c[3] : 0.0f;
        return ok;
    }
    static bool in_cube(const float* n) {
        return n[0] >= -1.0f && n[0] <= 1.0f && n[1] >= -1.0f && n[1] <= 1.0f && n[2] >= -1.0f && n[2] <= 1.0f;
    }
    float len2(const float* A, const float* B) const {
        float ca[4], cb[4], na[3], nb[3];
        clip(A, ca);
        clip(B, cb);
        const bool a_ok = ndc(ca, na) && in_cube(na);
        const bool b_ok = ndc(cb, nb) && in_cube(nb);
        if (!a_ok || !b_ok) return 0.0f;
        const float dx = (na[0] - nb[0]) * aspect;
        const float dy = na[1] - nb[1];
        return dx * dx + dy * dy;
    }
    void split_point(const float* A, const float* B, float* M) const {
        float ca[4], cb[4], na[3], nb[3];
        clip(A, ca);
        clip(B, cb);
        ndc(ca, na);
        ndc(cb, nb);
        const float mx = 0.5f * (na[0] + nb[0]);
        const float my = 0.5f * (na[1] + nb[1]);
        // clip(A + t (B - A)).xy - (mx, my) clip(...).w = r + t s; least squares in t
        const float dx = cb[0] - ca[0], dy = cb[1] - ca[1], dw = cb[3] - ca[3];
        const float rx = ca[0] - mx * ca[3];
        const float ry = ca[1] - my * ca[3];
        const float sx = dx - mx * dw;
        const float sy = dy - my * dw;
        const float ss = sx * sx + sy * sy;
        const float tb = (ss != 0.0f) ? -(rx * sx + ry * sy) / ss : 0.5f;
        const float ta = 1.0f - tb;
        for (int k = 0; k < 3; ++k) M[k] = ta * A[k] + tb * B[k];
    }
};

template <class Metric>
class Bisector {
public:
    Bisector(ds_soup_split& s, const Metric& metric) : s_(s), metric_(metric) {}

    // Split until `budget` splits are done (budget < 0: no cap) or the longest
    // queued edge is shorter than `limit` (limit <= 0: no threshold).
    // `eligible` (may be null) selects the triangles whose edges are queued
    // at the start; triangles created by splits are always eligible.
    void run(int budget, float limit, const int* eligible) {
        if (budget == 0) return;
        if (budget < 0 && !(limit > 0.0f)) return;   // unbounded needs a threshold
        const int T0 = n_tris();
        for (int t = 0; t < T0; ++t)
            if (eligible == nullptr || eligible[t] > 0) queue_edges(t);
        const float limit2 = (limit <= 0.0f) ? -1.0f : limit * limit;
        int done = 0;
        while ((budget < 0 || done < budget) && !q_.empty()) {
            const Candidate top = q_.top();
            q_.pop();
            if (top.epoch != s_.epoch[top.tri]) continue;
            if (limit2 >= 0.0f && top.len2 < limit2) break;
            const int fresh = bisect(top.tri, top.edge);
            queue_edges(top.tri);
            queue_edges(fresh);
            ++done;
        }
    }

private:
    int n_tris() const { return static_cast<int>(s_.corner.size() / 3); }
    const float* at(int v) const { return &s_.xyz[3 * static_cast<size_t>(v)]; }

    int push_vertex(const float* P) {
        const int v = static_cast<int>(s_.xyz.size() / 3);
        s_.xyz.insert(s_.xyz.end(), P, P + 3);
        return v;
    }

    void queue_edges(int t) {
        const int* c = &s_.corner[3 * static_cast<size_t>(t)];
        for (int e = 0; e < 3; ++e) {
            const float l2 = metric_.len2(at(c[e]), at(c[(e + 1) % 3]));
            if (Metric::kSkipZero && !(l2 > 0.0f)) continue;
            q_.push(Candidate{l2, t, e, s_.epoch[t]});
        }
    }

    int bisect(int t, int e) {
        int* c = &s_.corner[3 * static_cast<size_t>(t)];
        const int a = c[e], b = c[(e + 1) % 3], o = c[(e + 2) % 3];
        float M[3];
        metric_.split_point(at(a), at(b), M);
        const int m1 = push_vertex(M);
        const int m2 = push_vertex(M);
        float O[3] = {at(o)[0], at(o)[1], at(o)[2]};
        const int o1 = push_vertex(O);
        c[0] = a;
        c[1] = m1;
        c[2] = o1;
        const int fresh = n_tris();
        const int tri_new[3] = {m2, b, o};
        s_.corner.insert(s_.corner.end(), tri_new, tri_new + 3);
        s_.source.push_back(s_.source[t]);
        s_.pristine[t] = 0;
        s_.pristine.push_back(0);
        s_.epoch[t] += 1;
        s_.epoch.push_back(s_.epoch[t]);
        return fresh;
    }

    ds_soup_split& s_;
    Metric metric_;
    EdgeQueue q_;
};

int load(const float* verts, int nv, const int* faces, int nt, ds_soup_split** out, ds_soup_split** made) {
    if (out == nullptr) {
        return remesh_fail(-1, "split: null result pointer");
    }
    *out = nullptr;
    if (nv < 0 || nt < 0 || (nv > 0 && verts == nullptr) || (nt > 0 && faces == nullptr)) {
        return remesh_fail(-1, "split: bad vertex or face arrays");
    }
    for (long long i = 0; i < 3LL * nt; ++i) {
        if (faces[i] < 0 || faces[i] >= nv) {
            const std::string msg = "split: face " + std::to_string(i / 3) + " indexes vertex " +
                                    std::to_string(faces[i]) + " outside [0, " + std::to_string(nv) + ")";
            return remesh_fail(-1, msg.c_str());
        }
    }
    auto* s = new ds_soup_split;
    s->xyz.assign(verts, verts + 3 * static_cast<size_t>(nv));
    s->corner.assign(faces, faces + 3 * static_cast<size_t>(nt));
    s->source.resize(nt);
    for (int t = 0; t < nt; ++t) s->source[t] = t;
    s->pristine.assign(nt, 1);
    s->epoch.assign(nt, 0);
    *made = s;
    return 0;
}

}  // namespace
}  // namespace ds

extern "C" {

int ds_split_soup(const float* verts, int nv, const int* faces, int nt, int num_splits, float tau,
                  ds_soup_split** out) {
    ds_soup_split* s = nullptr;
    if (ds::load(verts, nv, faces, nt, out, &s) != 0) return -1;
    try {
        ds::Bisector<ds::WorldMetric>(*s, ds::WorldMetric{}).run(num_splits, tau, nullptr);
    } catch (const std::exception& ex) {
        delete s;
        return ds::remesh_fail(-3, (std::string("split: ") + ex.what()).c_str());
    }
    *out = s;
    return 0;
}

int ds_split_soup_clip(const float* mvp, const float* verts, int nv, const int* faces, int nt,
                       const int* valid_faces, int num_splits, float tau_ratio, float aspect_wh,
                       ds_soup_split** out) {
    if (mvp == nullptr) {
        if (out) *out = nullptr;
        return ds::remesh_fail(-1, "split_clip: null mvp");
    }
    ds_soup_split* s = nullptr;
    if (ds::load(verts, nv, faces, nt, out, &s) != 0) return -1;
    ds::ScreenMetric metric;
    std::memcpy(metric.m, mvp, sizeof(metric.m));
    metric.aspect = aspect_wh;
    try {
        // the threshold is in image heights; NDC spans two of them
        ds::Bisector<ds::ScreenMetric>(*s, metric).run(num_splits, 2.0f * tau_ratio, valid_faces);
    } catch (const std::exception& ex) {
        delete s;
        return ds::remesh_fail(-3, (std::string("split_clip: ") + ex.what()).c_str());
    }
    *out = s;
    return 0;
}

int ds_split_soup_counts(const ds_soup_split* s, int* nv, int* nt) {
    if (s == nullptr || nv == nullptr || nt == nullptr) return ds::remesh_fail(-1, "split_counts: null argument");
    *nv = static_cast<int>(s->xyz.size() / 3);
    *nt = static_cast<int>(s->corner.size() / 3);
    return 0;
}

int ds_split_soup_export(const ds_soup_split* s, float* verts, int* faces, int* face_map, int* same_as_original) {
    if (s == nullptr) return ds::remesh_fail(-1, "split_export: null result");
    if (verts) std::memcpy(verts, s->xyz.data(), s->xyz.size() * sizeof(float));
    if (faces) std::memcpy(faces, s->corner.data(), s->corner.size() * sizeof(int));
    if (face_map) std::memcpy(face_map, s->source.data(), s->source.size() * sizeof(int));
    if (same_as_original)
        for (size_t t = 0; t < s->pristine.size(); ++t) same_as_original[t] = s->pristine[t] ? 1 : 0;
    return 0;
}

void ds_split_soup_free(ds_soup_split* s) { delete s; }

}  // extern "C"
