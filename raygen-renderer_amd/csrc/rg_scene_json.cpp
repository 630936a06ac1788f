// rg_scene_json.cpp — host-side scene ingest behind rg_scene_load_json:
// the reference's scene JSON (SceneJsonLoader, sceneloader.cpp:101-165,
// 202-334, 505-783, 847-882), its primitive meshes (polygons.cpp:13-180),
// the binary .mesh format (meshloader.cpp:25-137), Radiance .hdr textures for
// the environment (texture.cpp:74-151) and the flattening of the object tree
// into render-space triangles (RayRenderer::transformScene / transformObject,
// rayrenderer.cpp:157-363; SceneTransformStack, renderer.cpp:24-73).
//
// Transform convention: a point is scaled, rotated (Euler angles in degrees,
// about X, then Y, then Z) and translated by its object, then by each
// parent; the camera's inverse (translate by -location, rotate by -angle)
// maps world to render space, where the eye sits at the origin looking down
// -Z.  The reference delegates these to its ugm Matrix4, which is not
// vendored, so the Euler order is the one documented here.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rg_internal.h"
#include "rg_json.h"

namespace rg {

// ---------------------------------------------------------------------------
// affine transforms (double precision, rounded to binary32 on output)
// ---------------------------------------------------------------------------
struct Affine {
    double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double t[3] = {0, 0, 0};

    // this ∘ o : apply o first, then this
    Affine then_after(const Affine &o) const {
        Affine r;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) r.m[i][j] = m[i][0] * o.m[0][j] + m[i][1] * o.m[1][j] + m[i][2] * o.m[2][j];
            r.t[i] = m[i][0] * o.t[0] + m[i][1] * o.t[1] + m[i][2] * o.t[2] + t[i];
        }
        return r;
    }
    void apply(const float p[3], double out[3]) const {
        for (int i = 0; i < 3; i++) out[i] = m[i][0] * p[0] + m[i][1] * p[1] + m[i][2] * p[2] + t[i];
    }
    // normal transform: inverse transpose of the linear part
    bool normal_matrix(double n[3][3]) const {
        const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
                     h = m[2][1], k = m[2][2];
        const double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
        if (det == 0.0) return false;
        const double inv[3][3] = {{(e * k - f * h) / det, (c * h - b * k) / det, (b * f - c * e) / det},
                                  {(f * g - d * k) / det, (a * k - c * g) / det, (c * d - a * f) / det},
                                  {(d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det}};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) n[i][j] = inv[j][i];
        return true;
    }
};

static Affine translation(const double v[3]) {
    Affine a;
    for (int i = 0; i < 3; i++) a.t[i] = v[i];
    return a;
}
static Affine scaling(const double v[3]) {
    Affine a;
    for (int i = 0; i < 3; i++) a.m[i][i] = v[i];
    return a;
}
static Affine rot_axis(int axis, double deg) {
    const double r = deg * M_PI / 180.0, c = std::cos(r), s = std::sin(r);
    Affine a;
    const int i = (axis + 1) % 3, j = (axis + 2) % 3;
    a.m[i][i] = c;
    a.m[i][j] = -s;
    a.m[j][i] = s;
    a.m[j][j] = c;
    return a;
}
static Affine euler(const double deg[3]) {
    // X first, then Y, then Z
    return rot_axis(2, deg[2]).then_after(rot_axis(1, deg[1]).then_after(rot_axis(0, deg[0])));
}

// ---------------------------------------------------------------------------
// meshes
// ---------------------------------------------------------------------------
struct MeshData {
    std::vector<float> v;   // triangle soup: 9 floats per triangle
    std::vector<float> n;   // 9 floats per triangle
    std::vector<float> uv;  // first UV set: 6 floats per triangle (zeros without UVs)
};

static const float kNoUv[2] = {0.0f, 0.0f};

static void push_tri(MeshData &m, const float *a, const float *b, const float *c, const float *na, const float *nb,
                     const float *nc, const float *ta = kNoUv, const float *tb = kNoUv, const float *tc = kNoUv) {
    m.v.insert(m.v.end(), a, a + 3);
    m.v.insert(m.v.end(), b, b + 3);
    m.v.insert(m.v.end(), c, c + 3);
    m.n.insert(m.n.end(), na, na + 3);
    m.n.insert(m.n.end(), nb, nb + 3);
    m.n.insert(m.n.end(), nc, nc + 3);
    m.uv.insert(m.uv.end(), {ta[0], ta[1], tb[0], tb[1], tc[0], tc[1]});
}

// PlaneMesh (polygons.cpp:13-59): unit plane in XZ, normal +Y
static MeshData plane_mesh() {
    MeshData m;
    const float v[6][3] = {{-0.5f, 0, -0.5f}, {-0.5f, 0, 0.5f}, {0.5f, 0, -0.5f},
                           {-0.5f, 0, 0.5f},  {0.5f, 0, 0.5f},  {0.5f, 0, -0.5f}};
    const float t[6][2] = {{0, 0}, {0, 1}, {1, 0}, {0, 1}, {1, 1}, {1, 0}};
    const float n[3] = {0, 1, 0};
    push_tri(m, v[0], v[1], v[2], n, n, n, t[0], t[1], t[2]);
    push_tri(m, v[3], v[4], v[5], n, n, n, t[3], t[4], t[5]);
    return m;
}

// CubeMesh (polygons.cpp:61-130): unit cube, flat face normals
static MeshData cube_mesh() {
    MeshData m;
    const float p[8][3] = {{-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f},
                           {-0.5f, -0.5f, 0.5f},  {0.5f, -0.5f, 0.5f},  {0.5f, 0.5f, 0.5f},  {-0.5f, 0.5f, 0.5f}};
    const int f[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 4, 7, 3}, {1, 5, 6, 2}, {3, 2, 6, 7}, {0, 1, 5, 4}};
    const float nn[6][3] = {{0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
    const float t[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 6; i++) {
        push_tri(m, p[f[i][0]], p[f[i][1]], p[f[i][2]], nn[i], nn[i], nn[i], t[0], t[1], t[2]);
        push_tri(m, p[f[i][0]], p[f[i][2]], p[f[i][3]], nn[i], nn[i], nn[i], t[0], t[2], t[3]);
    }
    return m;
}

// SphereMesh (polygons.cpp:132-180).  Its texcoords are never read: the
// constructor leaves uvCount at 0, so transformObject keeps zero UVs
// (rayrenderer.cpp:270-272; mesh.h:92).
static MeshData sphere_mesh(float radius, int stacks, int slices) {
    MeshData m;
    std::vector<float> pos, nrm;
    for (int i = 0; i <= stacks; ++i) {
        const float v = (float)i / stacks;
        const float phi = (float)(v * M_PI);
        for (int j = 0; j <= slices; ++j) {
            const float u = (float)j / slices;
            const float theta = (float)(u * 2.0f * M_PI);
            const float x = sinf(phi) * cosf(theta), y = cosf(phi), z = sinf(phi) * sinf(theta);
            pos.insert(pos.end(), {x * radius, y * radius, z * radius});
            nrm.insert(nrm.end(), {x, y, z});
        }
    }
    for (int i = 0; i < stacks; ++i)
        for (int j = 0; j < slices; ++j) {
            const int cur = i * (slices + 1) + j, next = (i + 1) * (slices + 1) + j;
            const int tri[2][3] = {{cur, next, cur + 1}, {cur + 1, next, next + 1}};
            for (auto &t : tri)
                push_tri(m, &pos[3 * t[0]], &pos[3 * t[1]], &pos[3 * t[2]], &nrm[3 * t[0]], &nrm[3 * t[1]],
                         &nrm[3 * t[2]]);
        }
    return m;
}

static bool read_file(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// Binary .mesh (meshloader.cpp:25-137): header {tag 'mesh', ver, flags,
// length}; meta; then vertices, normals, texcoords, tangents, colors,
// indexes.  Non-indexed meshes list 3 vertices per triangle.
static bool load_mesh_file(const std::string &path, MeshData &m, std::string &err) {
    std::string buf;
    if (!read_file(path, buf)) {
        err = "cannot read " + path;
        return false;
    }
    const uint8_t *p = (const uint8_t *)buf.data();
    const size_t sz = buf.size();
    auto rd = [&](size_t off, void *dst, size_t n) -> bool {
        if (off + n > sz) return false;
        std::memcpy(dst, p + off, n);
        return true;
    };
    uint32_t tag = 0, length = 0;
    uint16_t ver = 0, flags = 0;
    uint32_t vcount = 0, uvcount = 0, icount = 0;
    bool has_n = false, has_t = false, has_tb = false, has_c = false;
    size_t off = 0;
    if (!rd(0, &tag, 4)) {
        err = path + ": truncated";
        return false;
    }
    if (tag == 0x6873656du) {
        rd(4, &ver, 2);
        rd(6, &flags, 2);
        rd(8, &length, 4);
        uint32_t meta[5] = {0, 0, 0, 0, 0};
        if (!rd(12, meta, sizeof meta)) {
            err = path + ": truncated header";
            return false;
        }
        if (ver < 0x0102) {
            // vertexCount, normalCount, uvCount, texcoordCount, indexCount
            vcount = meta[0];
            uvcount = meta[2];
            icount = meta[4];
            has_n = meta[1] > 0;
            has_t = meta[3] > 0;
            has_tb = (flags & 0x2) != 0;
        } else {
            vcount = meta[0];
            uvcount = meta[1];
            icount = meta[2];
            has_n = (flags & 0x2) != 0;
            has_t = (flags & 0x4) != 0;
            has_tb = (flags & 0x10) != 0;
            has_c = ver >= 0x0103 && (flags & 0x20) != 0;
        }
        off = length;
    } else {
        uint32_t old[3];
        if (!rd(0, old, sizeof old)) {
            err = path + ": truncated";
            return false;
        }
        vcount = old[0];
        uvcount = 1;
        icount = 0;
        off = sizeof old;
    }
    if (!has_t) uvcount = 0;
    std::vector<float> verts(3 * (size_t)vcount), norms;
    if (!rd(off, verts.data(), verts.size() * 4)) {
        err = path + ": truncated vertices";
        return false;
    }
    off += verts.size() * 4;
    if (has_n) {
        norms.resize(3 * (size_t)vcount);
        if (!rd(off, norms.data(), norms.size() * 4)) {
            err = path + ": truncated normals";
            return false;
        }
        off += norms.size() * 4;
    }
    std::vector<float> uvs;
    if (has_t && uvcount > 0) {
        uvs.resize(2 * (size_t)vcount);  // the first of uvcount sets
        if (!rd(off, uvs.data(), uvs.size() * 4)) {
            err = path + ": truncated texcoords";
            return false;
        }
        off += 8 * (size_t)vcount * uvcount;
    }
    if (has_tb) off += 24 * (size_t)vcount;
    if (has_c) off += 12 * (size_t)vcount;
    std::vector<uint32_t> idx(icount);
    if (icount && !rd(off, idx.data(), idx.size() * 4)) {
        err = path + ": truncated indexes";
        return false;
    }
    const size_t ntri = icount > 0 ? icount / 3 : vcount / 3;
    const float zero[3] = {0, 0, 0};
    for (size_t k = 0; k < ntri; k++) {
        uint32_t i[3];
        for (int j = 0; j < 3; j++) i[j] = icount > 0 ? idx[3 * k + j] : (uint32_t)(3 * k + j);
        for (int j = 0; j < 3; j++)
            if (i[j] >= vcount) {
                err = path + ": index out of range";
                return false;
            }
        push_tri(m, &verts[3 * i[0]], &verts[3 * i[1]], &verts[3 * i[2]], has_n ? &norms[3 * i[0]] : zero,
                 has_n ? &norms[3 * i[1]] : zero, has_n ? &norms[3 * i[2]] : zero,
                 uvs.empty() ? kNoUv : &uvs[2 * i[0]], uvs.empty() ? kNoUv : &uvs[2 * i[1]],
                 uvs.empty() ? kNoUv : &uvs[2 * i[2]]);
    }
    return true;
}

// Face normals of a mesh read without normals, as Mesh::calcNormals
// computes them (mesh.cpp:531-550): normalize(cross(v1 - v3, v1 - v2)).
static void face_normals(MeshData &m) {
    const size_t nt = m.v.size() / 9;
    m.n.assign(m.v.size(), 0.0f);
    for (size_t t = 0; t < nt; t++) {
        const float *v = &m.v[9 * t];
        const float a[3] = {v[0] - v[6], v[1] - v[7], v[2] - v[8]};
        const float b[3] = {v[0] - v[3], v[1] - v[4], v[2] - v[5]};
        const float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        if (!(l > 0.0f)) continue;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) m.n[9 * t + 3 * k + j] = c[j] / l;
    }
}

// ---------------------------------------------------------------------------
// Wavefront OBJ: ObjFileReader::read (objreader.cpp:93-289), readSurfaceLine
// (:291-400), finalizeObject (:402-432), readMaterialLibrary (:490-590)
// ---------------------------------------------------------------------------
struct ObjMtl {
    std::string name, texture;
    float kd[3] = {0, 0, 0};
    float transparency = 0.0f;
};

struct ObjPart {
    std::string name, mat_name;
    int mtl = -1;  // index into the library, -1 when usemtl named none of it
    MeshData mesh;
};

// tag followed by a space (ObjFileReader::isLine)
static bool is_tag(const std::string &line, const char *tag) {
    const size_t n = std::strlen(tag);
    return line.size() > n && line.compare(0, n, tag) == 0 && line[n] == ' ';
}

static void read_mtl(const std::string &path, std::vector<ObjMtl> &out, std::vector<std::string> &warnings) {
    std::ifstream f(path);
    if (!f) {
        warnings.push_back("material library not read: " + path);
        return;
    }
    ObjMtl mat;
    bool has = false;
    std::string line;
    while (std::getline(f, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
        if (is_tag(line, "newmtl")) {
            if (has) out.push_back(mat);
            // ObjMaterial::reset clears name, texture and colours but keeps
            // the transparency of the previous entry
            const float keep = mat.transparency;
            mat = ObjMtl();
            mat.transparency = keep;
            mat.name = line.substr(7);
            has = true;
        } else if (is_tag(line, "Kd")) {
            float c[3];
            if (std::sscanf(line.c_str() + 3, "%f %f %f", &c[0], &c[1], &c[2]) >= 1)
                for (int k = 0; k < 3; k++) mat.kd[k] = c[k];
        } else if (is_tag(line, "Tf")) {
            float c[3];
            if (std::sscanf(line.c_str() + 3, "%f %f %f", &c[0], &c[1], &c[2]) >= 1)
                mat.transparency = 1.0f - (c[0] + c[1] + c[2]) / 3.0f;
        } else if (is_tag(line, "map_Kd")) {
            mat.texture = line.substr(7);
        }
    }
    if (has) out.push_back(mat);
}

// Reads `path` into its objects (fan-triangulated faces, 1-based indices,
// negative ones counted from the end as the OBJ format defines them).
// Objects whose faces mix texcoord or normal layouts are dropped, as the
// reference drops its error objects.
static bool load_obj(const std::string &path, std::vector<ObjPart> &parts, std::vector<ObjMtl> &mtls,
                     std::vector<std::string> &warnings, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot read " + path;
        return false;
    }
    const size_t slash = path.rfind('/');
    const std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    std::vector<float> vs, ns, ts;
    std::vector<ObjPart> done;
    ObjPart cur;
    bool cur_n = false, cur_t = false, first_face = true, cur_error = false, head = true, mm = false;
    int line_no = 0;
    auto finalize = [&]() {
        if (cur_error) warnings.push_back(path + ": object '" + cur.name + "' has inconsistent faces, skipped");
        else if (!cur.mesh.v.empty()) {
            if (!cur_n) face_normals(cur.mesh);
            done.push_back(std::move(cur));
        }
        cur = ObjPart();
        cur_n = cur_t = cur_error = false;
        first_face = true;
    };
    auto index = [](long v, size_t n) -> long { return v > 0 ? v - 1 : (v < 0 ? (long)n + v : -1); };
    std::string line;
    while (std::getline(f, line)) {
        line_no++;
        while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
        if (head && !line.empty() && line[0] == '#') {
            if (line.find("3ds Max") != std::string::npos ||
                line.find("uses millimeters as units") != std::string::npos) {
                mm = true;  // millimetre exports are scaled to metres
                head = false;
            }
            continue;
        }
        head = false;
        if (is_tag(line, "v")) {
            float x = 0, y = 0, z = 0;
            std::sscanf(line.c_str() + 2, "%f %f %f", &x, &y, &z);
            if (mm) {
                x *= 0.01f;
                y *= 0.01f;
                z *= 0.01f;
            }
            vs.insert(vs.end(), {x, y, z});
        } else if (is_tag(line, "vn")) {
            float x = 0, y = 0, z = 0;
            std::sscanf(line.c_str() + 3, "%f %f %f", &x, &y, &z);
            ns.insert(ns.end(), {x, y, z});
        } else if (is_tag(line, "vt")) {
            float u = 0, v = 0;
            std::sscanf(line.c_str() + 3, "%f %f", &u, &v);
            ts.insert(ts.end(), {u, 1.0f - v});  // OBJ's V origin is the image bottom
        } else if (is_tag(line, "f")) {
            if (cur_error) continue;
            std::vector<long> vi, ni, ti;
            bool has_t = false, has_n = false, ok = true;
            const char *p = line.c_str() + 2;
            while (*p && ok) {
                while (*p == ' ' || *p == '\t') p++;
                if (!*p) break;
                char *e = nullptr;
                const long v = std::strtol(p, &e, 10);
                if (e == p) { ok = false; break; }
                vi.push_back(index(v, vs.size() / 3));
                p = e;
                long t = 0, n = 0;
                bool gt = false, gn = false;
                if (*p == '/') {
                    p++;
                    if (*p != '/' && *p && *p != ' ') {
                        t = std::strtol(p, &e, 10);
                        if (e == p) { ok = false; break; }
                        gt = true;
                        p = e;
                    }
                    if (*p == '/') {
                        p++;
                        if (*p && *p != ' ') {
                            n = std::strtol(p, &e, 10);
                            if (e == p) { ok = false; break; }
                            gn = true;
                            p = e;
                        }
                    }
                }
                if (gt) ti.push_back(index(t, ts.size() / 2));
                // every face of an object carries the same attribute set
                if (gt ? (!cur_t && !first_face) : cur_t) ok = false;
                if (gn ? (!cur_n && !first_face) : cur_n) ok = false;
                if (gt) cur_t = has_t = true;
                if (gn) {
                    cur_n = has_n = true;
                    ni.push_back(index(n, ns.size() / 3));
                }
            }
            if (ok && vi.size() < 3) ok = false;
            for (long v : vi)
                if (v < 0 || (size_t)v >= vs.size() / 3) ok = false;
            for (long n : ni)
                if (n < 0 || (size_t)n >= ns.size() / 3) ok = false;
            if (has_n && ni.size() != vi.size()) ok = false;
            for (long t : ti)
                if (t < 0 || (size_t)t >= ts.size() / 2) ok = false;
            if (has_t && ti.size() != vi.size()) ok = false;
            if (!ok) {
                warnings.push_back(path + ": invalid face at line " + std::to_string(line_no));
                cur_error = true;
                continue;
            }
            const float zero[3] = {0, 0, 0};
            for (size_t i = 1; i + 1 < vi.size(); i++) {
                const float *a = &vs[3 * vi[0]], *b = &vs[3 * vi[i]], *c = &vs[3 * vi[i + 1]];
                push_tri(cur.mesh, a, b, c, has_n ? &ns[3 * ni[0]] : zero, has_n ? &ns[3 * ni[i]] : zero,
                         has_n ? &ns[3 * ni[i + 1]] : zero, has_t ? &ts[2 * ti[0]] : kNoUv,
                         has_t ? &ts[2 * ti[i]] : kNoUv, has_t ? &ts[2 * ti[i + 1]] : kNoUv);
            }
            first_face = false;
        } else if (is_tag(line, "o")) {
            finalize();
            cur.name = line.substr(2);
        } else if (is_tag(line, "g")) {
            if (line.substr(2) == "default") continue;
            finalize();
            std::istringstream ss(line.substr(2));
            std::string g, name;
            while (ss >> g) name += (name.empty() ? "" : "_") + g;
            cur.name = name;
        } else if (is_tag(line, "mtllib")) {
            read_mtl(dir + line.substr(7), mtls, warnings);
        } else if (is_tag(line, "usemtl")) {
            cur.mat_name = line.substr(7);
            cur.mtl = -1;
            for (size_t k = 0; k < mtls.size(); k++)
                if (mtls[k].name == cur.mat_name) {
                    cur.mtl = (int)k;
                    break;
                }
        }
    }
    finalize();
    parts = std::move(done);
    return true;
}

bool decode_jpeg(const uint8_t *data, size_t size, uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgb,
                 std::string &err);
bool read_rgbe(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgb, std::string &err);

// Binary PPM (P6, maxval <= 255).
static bool read_ppm(const std::string &buf, uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgb) {
    std::istringstream ss(buf);
    std::string magic;
    int W = 0, H = 0, mx = 0;
    ss >> magic;
    auto skip = [&]() {
        while (ss >> std::ws && ss.peek() == '#') {
            std::string c;
            std::getline(ss, c);
        }
    };
    skip(); ss >> W; skip(); ss >> H; skip(); ss >> mx;
    if (magic != "P6" || W <= 0 || H <= 0 || mx <= 0 || mx > 255) return false;
    ss.get();
    const size_t at = (size_t)ss.tellg();
    if (at + 3 * (size_t)W * H > buf.size()) return false;
    rgb.assign(buf.begin() + (long)at, buf.begin() + (long)(at + 3 * (size_t)W * H));
    w = (uint32_t)W;
    h = (uint32_t)H;
    return true;
}

// A material texture as Texture::loadFromFile reads it (texture.cpp:177-195):
// Radiance .hdr as linear floats, other images as sRGB bytes / 255.  JPEG
// (our decoder) and binary PPM are read; other formats are reported.
static bool load_texture(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgb, bool &srgb,
                         std::string &err) {
    const size_t dot = path.rfind('.');
    std::string ext = dot == std::string::npos ? "" : path.substr(dot);
    for (char &ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".hdr") {
        srgb = false;
        return read_rgbe(path, w, h, rgb, err);
    }
    std::string buf;
    if (!read_file(path, buf)) {
        err = "cannot read " + path;
        return false;
    }
    std::vector<uint8_t> px;
    bool ok = false;
    if (ext == ".jpg" || ext == ".jpeg") ok = decode_jpeg((const uint8_t *)buf.data(), buf.size(), w, h, px, err);
    else if (ext == ".ppm") {
        ok = read_ppm(buf, w, h, px);
        if (!ok) err = path + ": not a binary PPM";
    } else {
        err = "texture format not read on this path: " + path;
    }
    if (!ok) return false;
    srgb = true;
    rgb.resize(px.size());
    for (size_t i = 0; i < px.size(); i++) rgb[i] = (float)px[i] / 255.0f;
    return true;
}

// Radiance RGBE (.hdr) reader: flat and new-style run-length scanlines.
bool read_rgbe(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgb, std::string &err) {
    std::string buf;
    if (!read_file(path, buf)) {
        err = "cannot read " + path;
        return false;
    }
    size_t pos = 0;
    auto line = [&]() {
        const size_t e = buf.find('\n', pos);
        std::string l = buf.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
        pos = e == std::string::npos ? buf.size() : e + 1;
        return l;
    };
    std::string l = line();
    if (l.rfind("#?", 0) != 0) {
        err = path + ": not a Radiance HDR file";
        return false;
    }
    for (;;) {
        if (pos >= buf.size()) {
            err = path + ": missing resolution line";
            return false;
        }
        l = line();
        if (l.empty()) break;
    }
    l = line();
    int H = 0, W = 0;
    char ys[3] = {0}, xs[3] = {0};
    if (std::sscanf(l.c_str(), "%2s %d %2s %d", ys, &H, xs, &W) != 4 || std::strcmp(ys, "-Y") || std::strcmp(xs, "+X") ||
        W <= 0 || H <= 0) {
        err = path + ": unsupported resolution line '" + l + "'";
        return false;
    }
    w = (uint32_t)W;
    h = (uint32_t)H;
    rgb.assign((size_t)W * H * 3, 0.0f);
    std::vector<uint8_t> scan((size_t)W * 4);
    const uint8_t *p = (const uint8_t *)buf.data();
    for (int y = 0; y < H; y++) {
        if (pos + 4 > buf.size()) {
            err = path + ": truncated";
            return false;
        }
        if (W >= 8 && W < 32768 && p[pos] == 2 && p[pos + 1] == 2 && ((p[pos + 2] << 8) | p[pos + 3]) == W) {
            pos += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < W) {
                    if (pos >= buf.size()) {
                        err = path + ": truncated scanline";
                        return false;
                    }
                    int cnt = p[pos++];
                    if (cnt > 128) {
                        cnt -= 128;
                        if (x + cnt > W || pos >= buf.size()) {
                            err = path + ": bad run";
                            return false;
                        }
                        const uint8_t v = p[pos++];
                        for (int k = 0; k < cnt; k++) scan[4 * (x++) + c] = v;
                    } else {
                        if (cnt == 0 || x + cnt > W || pos + cnt > buf.size()) {
                            err = path + ": bad literal";
                            return false;
                        }
                        for (int k = 0; k < cnt; k++) scan[4 * (x++) + c] = p[pos++];
                    }
                }
            }
        } else {
            if (pos + 4 * (size_t)W > buf.size()) {
                err = path + ": truncated flat scanline";
                return false;
            }
            std::memcpy(scan.data(), p + pos, 4 * (size_t)W);
            pos += 4 * (size_t)W;
        }
        for (int x = 0; x < W; x++) {
            const uint8_t *q = &scan[4 * x];
            float f = 0.0f;
            if (q[3]) f = std::ldexp(1.0f, (int)q[3] - (128 + 8));
            float *o = &rgb[3 * ((size_t)y * W + x)];
            o[0] = q[0] * f;
            o[1] = q[1] * f;
            o[2] = q[2] * f;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// the loader
// ---------------------------------------------------------------------------
namespace {

// A scene object's Material as far as this path reads it: the shading
// fields, the albedo texture path and texTiling (material.h:20-70).
struct Mat {
    rg_object o;
    std::string tex;
    float tiling[2] = {1.0f, 1.0f};
    float spot_range = 0.0f;  // Material::spotRange (material.h:52), read by point lights
    Mat() { rg_object_default(&o); }
};

struct Loader {
    std::string base;
    std::vector<float> v9, n9, uv6;
    std::vector<uint32_t> obj_of_tri;
    std::vector<rg_object> objects;
    std::vector<std::string> obj_tex;      // texture path per object ("" none)
    std::vector<float> obj_tiling;         // 2 per object
    std::vector<rg_medium> media;
    std::vector<Affine> media_model;       // owning object's model matrix (identity: global)
    std::vector<rg_volume_light> volume_lights;  // pos: object location, world space until flatten
    int32_t global_medium = -1;
    std::vector<std::vector<std::pair<std::string, Mat>>> scopes;  // _materials per level
    bool has_camera = false;
    rg_camera camera{75.0f, 1.0f, 0.0f, 1.8f, 0, 0.0f, 50.0f, 0.0f};
    Affine view;  // world -> render
    std::string env_path;
    float env_intensity = 1.0f, env_rotation = 0.0f;
    std::vector<std::string> warnings;

    std::string resolve(const std::string &p) const {
        if (!p.empty() && p[0] == '/') return p;
        return base + p;
    }

    static bool vec3_of(const json::Value *v, double out[3]) {
        if (!v) return false;
        if (v->is_array() && v->arr.size() >= 3) {
            for (int k = 0; k < 3; k++) out[k] = v->arr[k].is_number() ? v->arr[k].num : 0.0;
            return true;
        }
        if (v->is_object()) {
            const char *names[3] = {"x", "y", "z"};
            for (int k = 0; k < 3; k++) {
                const json::Value *c = v->get(names[k]);
                out[k] = c && c->is_number() ? c->num : 0.0;
            }
            return true;
        }
        return false;
    }
    static void num(const json::Value &o, const char *key, float &dst) {
        const json::Value *v = o.get(key);
        if (v && v->is_number()) dst = (float)v->num;
    }
    static void inum(const json::Value &o, const char *key, int32_t &dst) {
        const json::Value *v = o.get(key);
        if (v && v->is_number()) dst = (int32_t)v->num;
    }
    static bool is_true(const json::Value &o, const char *key) {
        const json::Value *v = o.get(key);
        return v && v->kind == json::Value::Bool && v->b;
    }
    static void vec3f(const json::Value &o, const char *key, float dst[3]) {
        double v[3];
        if (vec3_of(o.get(key), v))
            for (int k = 0; k < 3; k++) dst[k] = (float)v[k];
    }

    // colour string: "#rrggbb", "#rrggbbaa" (and the reference's 3/4-digit forms)
    static bool parse_color(const std::string &s, float out[3]) {
        std::string h = s;
        if (!h.empty() && h[0] == '#') h = h.substr(1);
        if (h.size() != 3 && h.size() != 4 && h.size() != 6 && h.size() != 8) return false;
        unsigned long v = std::strtoul(h.c_str(), nullptr, 16);
        if (h.size() == 3 || h.size() == 4) v *= 2;  // sceneloader.cpp:181-186
        if (h.size() == 8 || h.size() == 4) v >>= 8;
        out[0] = (float)((v >> 16) & 0xFF) / 255.0f;
        out[1] = (float)((v >> 8) & 0xFF) / 255.0f;
        out[2] = (float)(v & 0xFF) / 255.0f;
        return true;
    }

    // SceneJsonLoader::readMaterial (sceneloader.cpp:101-165)
    void read_material(Mat &mt, const json::Value &o) {
        rg_object &m = mt.o;
        num(o, "emission", m.emission);
        num(o, "glossy", m.glossy);
        num(o, "roughness", m.roughness);
        num(o, "transparency", m.transparency);
        num(o, "refraction", m.refraction);
        num(o, "refractionRatio", m.refraction_ratio);
        num(o, "metallic", m.metallic);
        num(o, "anisotropy", m.anisotropy);
        num(o, "anisoRotation", m.aniso_rotation);
        num(o, "chromaDispersion", m.chroma_dispersion);
        const json::Value *c = o.get("color");
        if (c && c->is_string()) parse_color(c->str, m.color);
        else if (c && c->is_array() && c->arr.size() >= 3)
            for (int k = 0; k < 3; k++)
                if (c->arr[k].is_number()) m.color[k] = (float)c->arr[k].num;
        const json::Value *t = o.get("tex");
        if (t && t->is_string() && !t->str.empty()) mt.tex = resolve(t->str);
        if (const json::Value *tt = o.get("texTiling")) {
            if (tt->is_array() && tt->arr.size() >= 2) {
                for (int k = 0; k < 2; k++)
                    if (tt->arr[k].is_number()) mt.tiling[k] = (float)tt->arr[k].num;
            } else if (tt->is_object()) {
                num(*tt, "x", mt.tiling[0]);
                num(*tt, "y", mt.tiling[1]);
            }
        }
        num(o, "spotRange", mt.spot_range);
        // read by the reference but outside the -d 5 path kept here
        if (o.get("normalmap")) warnings.push_back("normal map ignored");
    }

    const Mat *find_material(const std::string &name) const {
        for (auto s = scopes.rbegin(); s != scopes.rend(); ++s)
            for (auto it = s->rbegin(); it != s->rend(); ++it)
                if (it->first == name) return &it->second;
        return nullptr;
    }

    // SceneJsonLoader::readMedium (sceneloader.cpp:202-334), homogeneous part
    // SceneJsonLoader::readMedium (sceneloader.cpp:202-334); `model` is the
    // owning object's model matrix (identity for the global medium), used
    // when the cone or path follows its object (medium.cpp:59-104)
    int read_medium(const json::Value &o, const Affine &model) {
        rg_medium m;
        rg_medium_default(&m);
        vec3f(o, "sigma_a", m.sigma_a);
        vec3f(o, "sigma_s", m.sigma_s);
        vec3f(o, "emission", m.sigma_e);
        num(o, "g", m.g);
        num(o, "density", m.density);
        const json::Value *mode = o.get("emissionMode");
        if (mode && mode->is_string() && mode->str == "cone") m.emission_mode = 1;
        else if (mode && mode->is_string() && mode->str == "path") m.emission_mode = 2;
        vec3f(o, "coneAxis", m.cone_axis);
        vec3f(o, "coneOrigin", m.cone_origin);
        num(o, "coneLength", m.cone_length);
        num(o, "coneRadius", m.cone_radius);
        vec3f(o, "coneInner", m.cone_inner);
        vec3f(o, "coneOuter", m.cone_outer);
        num(o, "coneIntensity", m.cone_intensity);
        num(o, "conePeakAxial", m.cone_peak_axial);
        num(o, "conePeakSharpness", m.cone_peak_sharpness);
        inum(o, "coneEmissionSamples", m.cone_samples);
        const bool cone_follow = is_true(o, "coneFollowObject");
        vec3f(o, "pathInner", m.path_inner);
        vec3f(o, "pathOuter", m.path_outer);
        num(o, "pathIntensity", m.path_intensity);
        num(o, "pathFalloffPower", m.path_falloff_power);
        inum(o, "pathEmissionSamples", m.path_samples);
        const bool path_follow = is_true(o, "pathFollowObject");
        if (const json::Value *pts = o.get("pathPoints")) {
            if (pts->is_array()) {
                for (const json::Value &e : pts->arr) {
                    if (!e.is_object()) continue;
                    if (m.n_path == RG_MAX_PATH_POINTS) {
                        warnings.push_back("medium pathPoints beyond 16 ignored");
                        break;
                    }
                    rg_path_point &q = m.path[m.n_path++];
                    q.p[0] = q.p[1] = q.p[2] = 0.0f;
                    q.radius = 0.1f;
                    q.t = 0.0f;
                    vec3f(e, "p", q.p);
                    num(e, "radius", q.radius);
                    num(e, "t", q.t);
                }
            }
        }
        const json::Value *df = o.get("densityField");
        if (df && df->is_string() && df->str == "fbm") m.density_field = 1;
        num(o, "noiseFrequency", m.noise_frequency);
        inum(o, "noiseOctaves", m.noise_octaves);
        num(o, "noiseGain", m.noise_gain);
        num(o, "noiseLacunarity", m.noise_lacunarity);
        num(o, "noiseAmplitude", m.noise_amplitude);
        num(o, "noiseBias", m.noise_bias);
        vec3f(o, "noiseOffset", m.noise_offset);
        if (o.get("heatHaze")) m.heat_haze = is_true(o, "heatHaze") ? 1 : 0;
        num(o, "iorAmplitude", m.ior_amplitude);
        num(o, "iorFrequency", m.ior_frequency);
        inum(o, "iorOctaves", m.ior_octaves);
        num(o, "iorGain", m.ior_gain);
        num(o, "iorLacunarity", m.ior_lacunarity);
        inum(o, "iorMarchSteps", m.ior_march_steps);
        vec3f(o, "iorOffset", m.ior_offset);
        num(o, "iorFalloff", m.ior_falloff);
        // prepare (medium.cpp:18-57): the cone axis normalised, -Z when degenerate
        const float al2 = m.cone_axis[0] * m.cone_axis[0] + m.cone_axis[1] * m.cone_axis[1] +
                          m.cone_axis[2] * m.cone_axis[2];
        if (al2 > 1e-12f) {
            const float inv = 1.0f / sqrtf(al2);
            for (float &x : m.cone_axis) x *= inv;
        } else {
            m.cone_axis[0] = m.cone_axis[1] = 0.0f;
            m.cone_axis[2] = -1.0f;
        }
        // an inactive medium is dropped (isActive, medium.h:214-223;
        // sceneloader.cpp:326-332)
        const float d = fmaxf(0.0f, m.density);
        float st[3];
        bool emits = false;
        for (int k = 0; k < 3; k++) {
            st[k] = m.sigma_a[k] * d + m.sigma_s[k] * d;
            emits = emits || m.sigma_e[k] * d != 0.0f;
        }
        const float hero = (st[0] + st[1] + st[2]) * (1.0f / 3.0f);
        const bool active = hero > 0.0f || (m.emission_mode == 1 && m.cone_intensity > 0.0f) ||
                            (m.emission_mode == 2 && m.path_intensity > 0.0f && m.n_path >= 2) ||
                            (m.heat_haze && m.ior_amplitude > 0.0f) || emits;
        if (!active) return -1;
        media.push_back(m);
        // the cone's / path's own frame: the object's, or world space
        media_model.push_back((m.emission_mode == 1 && cone_follow) || (m.emission_mode == 2 && path_follow)
                                  ? model : Affine());
        return (int)media.size() - 1;
    }

    bool is_camera_like(const std::string &key) const { return key == "mainCamera"; }

    void read_camera(const json::Value &o) {
        has_camera = true;
        num(o, "fieldOfView", camera.fov_deg);
        num(o, "depthOfField", camera.depth_of_field);
        num(o, "aperture", camera.aperture);
        float blades = (float)camera.aperture_blades;
        num(o, "apertureBlades", blades);
        camera.aperture_blades = (int32_t)blades;
        num(o, "apertureRotation", camera.aperture_rotation_deg);
        num(o, "exposure", camera.exposure);
        num(o, "viewFar", camera.view_far);
        if (o.get("focusOn")) warnings.push_back("camera focusOn ignored");
        double loc[3] = {0, 0, 0}, ang[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
        vec3_of(o.get("location"), loc);
        vec3_of(o.get("angle"), ang);
        vec3_of(o.get("scale"), sc);
        // Renderer::applyCameraTransform (renderer.cpp:24-28): rotate(-angle),
        // translate(-location), scale(1/scale)
        const double nl[3] = {-loc[0], -loc[1], -loc[2]};
        const double na[3] = {-ang[0], -ang[1], -ang[2]};
        const double is[3] = {1.0 / sc[0], 1.0 / sc[1], 1.0 / sc[2]};
        // inverse rotation of an X-then-Y-then-Z Euler rotation: Z, then Y, then X
        const Affine inv_rot = rot_axis(0, na[0]).then_after(rot_axis(1, na[1]).then_after(rot_axis(2, na[2])));
        view = inv_rot.then_after(translation(nl).then_after(scaling(is)));
    }

    std::string err;

    // A child of an object: a JSON object, or an object of an OBJ file with
    // its resolved material (SceneJsonLoader::readObjAsSceneObjects,
    // sceneloader.cpp:424-478), in the order the reference adds them.
    struct Child {
        const json::Value *json = nullptr;
        Mat mat;
        MeshData mesh;
    };

    // .obj as child objects; `parent_mat` is the loading object's material
    // at the point its "mesh" key is read
    bool read_obj(const std::string &path, const Mat &parent_mat, std::vector<Child> &out) {
        std::vector<ObjPart> parts;
        std::vector<ObjMtl> mtls;
        if (!load_obj(path, parts, mtls, warnings, err)) return false;
        for (ObjPart &p : parts) {
            Child c;
            // _materials override by name > .mtl defaults > the parent's mat
            const Mat *over = p.mat_name.empty() ? nullptr : find_material(p.mat_name);
            if (over) {
                c.mat = *over;
            } else if (p.mtl >= 0) {
                // applyObjMaterialDefaults (sceneloader.cpp:405-422)
                const ObjMtl &m = mtls[(size_t)p.mtl];
                c.mat = Mat();
                for (int k = 0; k < 3; k++) c.mat.o.color[k] = m.kd[k];
                c.mat.o.transparency = m.transparency;
                if (!m.texture.empty()) {
                    const size_t sl = path.rfind('/');
                    c.mat.tex = (sl == std::string::npos ? std::string() : path.substr(0, sl + 1)) + m.texture;
                }
            } else {
                c.mat = parent_mat;
                c.mat.o.interior_medium = -1;
            }
            c.mesh = std::move(p.mesh);
            out.push_back(std::move(c));
        }
        return true;
    }

    // mesh property of an object (sceneloader.cpp:618-641, 390-403)
    bool read_meshes(const json::Value &val, std::vector<MeshData> &out, const Mat &cur_mat,
                     std::vector<Child> &kids) {
        auto from_path = [&](const std::string &rel) -> bool {
            const std::string p = resolve(rel);
            const size_t dot = p.rfind('.');
            std::string ext = dot == std::string::npos ? "" : p.substr(dot);
            for (char &ch : ext) ch = (char)std::tolower((unsigned char)ch);
            if (ext == ".obj") return read_obj(p, cur_mat, kids);
            if (ext == ".fbx") {
                warnings.push_back("FBX import is outside this path, skipped: " + rel);
                return true;
            }
            MeshData m;
            if (!load_mesh_file(p, m, err)) return false;
            out.push_back(std::move(m));
            return true;
        };
        if (val.is_string()) return from_path(val.str);
        if (val.is_array()) {
            for (const json::Value &x : val.arr)
                if (x.is_string() && !from_path(x.str)) return false;
            return true;
        }
        if (val.is_object()) {
            const json::Value *t = val.get("type");
            const std::string type = t && t->is_string() ? t->str : "";
            if (type == "plane") out.push_back(plane_mesh());
            else if (type == "cube") out.push_back(cube_mesh());
            else if (type == "sphere") {
                float r = 1, st = 16, sl = 32;
                num(val, "radius", r);
                num(val, "stacks", st);
                num(val, "slices", sl);
                out.push_back(sphere_mesh(r, (int)st, (int)sl));
            } else if (!type.empty()) {
                warnings.push_back("primitive '" + type + "' is outside this path, skipped");
            }
        }
        return true;
    }

    // readSceneObject + transformObject, depth first in file order
    bool object(const json::Value &o, const Affine &parent, bool is_root) {
        scopes.emplace_back();
        if (const json::Value *mats = o.get("_materials"))
            if (mats->is_object())
                for (auto &p : mats->obj)
                    if (p.second.is_object()) {
                        Mat m;
                        read_material(m, p.second);
                        scopes.back().emplace_back(p.first, m);
                    }
        double loc[3] = {0, 0, 0}, ang[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
        vec3_of(o.get("location"), loc);
        vec3_of(o.get("angle"), ang);
        vec3_of(o.get("scale"), sc);
        const Affine model = parent.then_after(translation(loc).then_after(euler(ang).then_after(scaling(sc))));
        Mat mat;
        int32_t interior = -1;
        std::vector<MeshData> meshes;
        std::vector<Child> children;
        bool visible = true;
        for (auto &p : o.obj) {
            const std::string &key = p.first;
            const json::Value &val = p.second;
            if (key == "_materials" || key == "_bundles" || key == "location" || key == "angle" || key == "scale") {
            } else if (key == "_bundle") {
                warnings.push_back(".toba bundles are outside this path, skipped");
            } else if (key == "mesh") {
                if (!read_meshes(val, meshes, mat, children)) return false;
            } else if (key == "mat") {
                if (val.is_string()) {
                    const Mat *m = find_material(val.str);
                    if (m) mat = *m;
                    else warnings.push_back("unknown material '" + val.str + "'");
                } else if (val.is_object()) {
                    read_material(mat, val);
                }
            } else if (key == "visible" && val.kind == json::Value::Bool) {
                visible = val.b;
            } else if (key == "medium" && val.is_object()) {
                const int m = read_medium(val, model);
                if (is_root) {
                    global_medium = m;
                    const rg_medium *g = m >= 0 ? &media[(size_t)m] : nullptr;
                    if (g && g->heat_haze && g->ior_amplitude > 0.0f && g->ior_march_steps > 0) {
                        // tracePath recurses without end in a heat-haze global medium
                        warnings.push_back("heat-haze global medium dropped");
                        global_medium = -1;
                    }
                } else {
                    interior = m;
                }
                if (!is_root && m >= 0) register_volume(m, loc);
            } else if (key == "envmap") {
                if (val.is_string()) env_path = resolve(val.str);
                else if (val.is_object()) {
                    const json::Value *t = val.get("texture");
                    if (t && t->is_string() && !t->str.empty()) env_path = resolve(t->str);
                    if (val.get("cubemap")) warnings.push_back("cubemap environments are outside this path");
                    num(val, "intensity", env_intensity);
                    num(val, "rotation", env_rotation);
                }
            } else if (is_camera_like(key) && val.is_object()) {
                read_camera(val);
            } else if (val.is_object()) {
                Child c;
                c.json = &val;
                children.push_back(std::move(c));
            }
        }
        if (!visible) {
            hidden_volumes(o, model);
            scopes.pop_back();
            return true;
        }
        mat.o.interior_medium = interior;
        if (!meshes.empty()) {
            const uint32_t oid = add_object(mat);
            pending.push_back({model, oid, std::move(meshes)});
        } else if (mat.o.emission > 0.0f) {
            // an emissive object without meshes is a point light source
            // (rayrenderer.cpp:333-353)
            pending_lights.push_back({model, mat, {ang[0], ang[1], ang[2]}});
        }
        for (Child &c : children) {
            if (c.json) {
                if (!object(*c.json, model, false)) return false;
            } else {
                // an OBJ object: a child with the identity transform
                const uint32_t oid = add_object(c.mat);
                std::vector<MeshData> one;
                one.push_back(std::move(c.mesh));
                pending.push_back({model, oid, std::move(one)});
            }
        }
        scopes.pop_back();
        return true;
    }

    // emissive-volume registration (rayrenderer.cpp:180-206): a cone of
    // positive intensity or constant non-zero emission, at the object's own
    // location (baked through the view matrix in flatten)
    void register_volume(int m, const double loc[3]) {
        const rg_medium &q = media[(size_t)m];
        const float d = fmaxf(0.0f, q.density);
        const bool cone = q.emission_mode == 1 && q.cone_intensity > 0.0f;
        const bool cst = q.emission_mode == 0 &&
                         (q.sigma_e[0] * d != 0.0f || q.sigma_e[1] * d != 0.0f || q.sigma_e[2] * d != 0.0f);
        if (!cone && !cst) return;
        volume_lights.push_back({(uint32_t)m, {(float)loc[0], (float)loc[1], (float)loc[2]}});
    }

    // objects under an invisible one are not rendered, but their media are
    // still registered as emissive volumes, as the reference's registration
    // walks every object (rayrenderer.cpp:185-204)
    void hidden_volumes(const json::Value &o, const Affine &parent) {
        for (auto &p : o.obj) {
            if (!p.second.is_object() || p.first == "_materials" || p.first == "_bundles" || p.first == "mesh" ||
                p.first == "mat" || p.first == "medium" || p.first == "envmap" || is_camera_like(p.first))
                continue;
            const json::Value &c = p.second;
            double loc[3] = {0, 0, 0}, ang[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
            vec3_of(c.get("location"), loc);
            vec3_of(c.get("angle"), ang);
            vec3_of(c.get("scale"), sc);
            const Affine model = parent.then_after(translation(loc).then_after(euler(ang).then_after(scaling(sc))));
            if (const json::Value *md = c.get("medium"))
                if (md->is_object()) {
                    const int m = read_medium(*md, model);
                    if (m >= 0) register_volume(m, loc);
                }
            hidden_volumes(c, model);
        }
    }

    uint32_t add_object(const Mat &m) {
        objects.push_back(m.o);
        obj_tex.push_back(m.tex);
        obj_tiling.insert(obj_tiling.end(), {m.tiling[0], m.tiling[1]});
        return (uint32_t)objects.size() - 1;
    }

    struct Pending {
        Affine model;
        uint32_t obj;
        std::vector<MeshData> meshes;
    };
    std::vector<Pending> pending;
    struct PendingLight {
        Affine model;
        Mat mat;
        double angle[3];
    };
    std::vector<PendingLight> pending_lights;
    std::vector<rg_point_light> point_lights;

    // transformObject (rayrenderer.cpp:237-363): render-space vertices and
    // normalised normals, once the camera is known
    void flatten() {
        // HomogeneousMedium::bake (medium.cpp:59-104): a cone's origin and
        // axis, a path's points, into render space through the view matrix
        // (and the object's model matrix when they follow it); media of
        // other modes keep their authored cone fields
        for (size_t i = 0; i < media.size(); i++) {
            rg_medium &m = media[i];
            if (m.emission_mode != 1 && m.emission_mode != 2) continue;
            const Affine full = view.then_after(media_model[i]);
            double q[3];
            if (m.emission_mode == 1) {
                full.apply(m.cone_origin, q);
                double a[3];
                for (int c = 0; c < 3; c++)
                    a[c] = full.m[c][0] * m.cone_axis[0] + full.m[c][1] * m.cone_axis[1] + full.m[c][2] * m.cone_axis[2];
                for (int c = 0; c < 3; c++) m.cone_origin[c] = (float)q[c];
                const double l2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
                if (l2 > 1e-12)
                    for (int c = 0; c < 3; c++) m.cone_axis[c] = (float)(a[c] / std::sqrt(l2));
            } else {
                for (uint32_t k = 0; k < m.n_path; k++) {
                    full.apply(m.path[k].p, q);
                    for (int c = 0; c < 3; c++) m.path[k].p[c] = (float)q[c];
                }
            }
        }
        for (rg_volume_light &v : volume_lights) {
            double q[3];
            view.apply(v.pos, q);
            for (int c = 0; c < 3; c++) v.pos[c] = (float)q[c];
        }
        // point lights (rayrenderer.cpp:339-352): the origin through the
        // view-model matrix; the direction from the object's angle.x / .y
        // read through RADIAN_TO_DEGREE (x * 180 / pi) in binary32, then the
        // normal matrix (not normalised again, as there)
        for (PendingLight &pl : pending_lights) {
            const Affine vm = view.then_after(pl.model);
            rg_point_light L;
            std::memset(&L, 0, sizeof L);
            const float zero[3] = {0.0f, 0.0f, 0.0f};
            double q[3];
            vm.apply(zero, q);
            const float r2d = (float)(180.0 / M_PI);
            const float ax = (float)pl.angle[0] * r2d, ay = (float)pl.angle[1] * r2d;
            const float s1 = sinf(ax), c1 = cosf(ax), s2 = sinf(ay), c2 = cosf(ay);
            float d[3] = {c2 * s1, c2 * c1, s2};
            const float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const float inv = dl > 0.0f ? 1.0f / dl : 0.0f;
            for (float &x : d) x *= inv;
            double nm[3][3];
            const bool ok = vm.normal_matrix(nm);
            for (int c = 0; c < 3; c++) {
                L.pos[c] = (float)q[c];
                L.normal[c] = ok ? (float)(nm[c][0] * d[0] + nm[c][1] * d[1] + nm[c][2] * d[2]) : d[c];
                L.color[c] = pl.mat.o.color[c];
            }
            L.emission = pl.mat.o.emission;
            L.spot_range = pl.mat.spot_range;
            point_lights.push_back(L);
        }
        for (Pending &p : pending) {
            const Affine vm = view.then_after(p.model);
            double nm[3][3];
            const bool ok = vm.normal_matrix(nm);
            for (MeshData &m : p.meshes) {
                const size_t nt = m.v.size() / 9;
                for (size_t t = 0; t < nt; t++) {
                    for (int k = 0; k < 3; k++) {
                        double q[3];
                        vm.apply(&m.v[9 * t + 3 * k], q);
                        for (int c = 0; c < 3; c++) v9.push_back((float)q[c]);
                        const float *nn = &m.n[9 * t + 3 * k];
                        double r[3] = {nn[0], nn[1], nn[2]};
                        if (ok)
                            for (int c = 0; c < 3; c++) r[c] = nm[c][0] * nn[0] + nm[c][1] * nn[1] + nm[c][2] * nn[2];
                        const double l = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                        for (int c = 0; c < 3; c++) n9.push_back(l > 0 ? (float)(r[c] / l) : 0.0f);
                    }
                    obj_of_tri.push_back(p.obj);
                    for (int k = 0; k < 6; k++) uv6.push_back(t < m.uv.size() / 6 ? m.uv[6 * t + k] : 0.0f);
                }
            }
        }
    }
};

}  // namespace

}  // namespace rg

using namespace rg;

extern "C" int rg_scene_load_json(rg_ctx *ctx, const char *path, rg_scene **out) {
    if (!ctx || !path || !out) return set_error(RG_E_INVALID, "null argument");
    *out = nullptr;
    std::string text;
    if (!read_file(path, text)) return set_error(RG_E_IO, std::string("cannot read ") + path);
    if (text.empty()) return set_error(RG_E_IO, "scene file is empty");
    Loader L;
    const std::string p(path);
    const size_t slash = p.rfind('/');
    L.base = slash == std::string::npos ? "" : p.substr(0, slash + 1);
    try {
        json::Parser parser(text);
        const json::Value root = parser.parse();
        if (!root.is_object()) return set_error(RG_E_IO, "scene root is not an object");
        Affine ident;
        if (!L.object(root, ident, true)) return set_error(RG_E_IO, L.err);
    } catch (const std::exception &e) {
        return set_error(RG_E_IO, e.what());
    }
    L.flatten();
    rg_scene *sc = nullptr;
    if (int rc = rg_scene_create(ctx, L.v9.data(), (uint32_t)L.obj_of_tri.size(), &sc)) return rc;
    int rc = rg_scene_set_media(sc, L.media.data(), (uint32_t)L.media.size(), L.global_medium);
    if (!rc) rc = rg_scene_set_shading(sc, L.n9.data(), L.obj_of_tri.data(), L.objects.data(), (uint32_t)L.objects.size());
    if (!rc) rc = rg_scene_set_camera(sc, &L.camera);
    if (!rc && !L.point_lights.empty())
        rc = rg_scene_set_point_lights(sc, L.point_lights.data(), (uint32_t)L.point_lights.size());
    if (!rc && !L.volume_lights.empty())
        rc = rg_scene_set_volume_lights(sc, L.volume_lights.data(), (uint32_t)L.volume_lights.size());
    if (!rc && !L.env_path.empty()) {
        const size_t dot = L.env_path.rfind('.');
        const std::string ext = dot == std::string::npos ? "" : L.env_path.substr(dot);
        if (ext == ".hdr" || ext == ".HDR") {
            uint32_t w = 0, h = 0;
            std::vector<float> rgb;
            std::string err;
            if (read_rgbe(L.env_path, w, h, rgb, err)) {
                rg_env env{w, h, rgb.data(), L.env_intensity, L.env_rotation};
                rc = rg_scene_set_env(sc, &env);
            } else {
                L.warnings.push_back("environment not loaded: " + err);
            }
        } else {
            L.warnings.push_back("only .hdr environments are read on this path: " + L.env_path);
        }
    }
    if (!rc) {
        // material textures, one per distinct path; a texture that cannot be
        // read is dropped and its objects keep their flat colour, as
        // SceneResourcePool::getTexture drops a failed decode (scene.cpp:656-664)
        std::vector<std::string> paths;
        std::vector<std::vector<float>> texels;
        std::vector<rg_texture> tex;
        std::vector<int32_t> tex_of_obj(L.objects.size(), -1);
        for (size_t o = 0; o < L.objects.size(); o++) {
            const std::string &tp = L.obj_tex[o];
            if (tp.empty()) continue;
            size_t k = 0;
            while (k < paths.size() && paths[k] != tp) k++;
            if (k == paths.size()) {
                paths.push_back(tp);
                texels.emplace_back();
                uint32_t w = 0, h = 0;
                bool srgb = true;
                std::string err;
                if (load_texture(tp, w, h, texels.back(), srgb, err)) {
                    tex.push_back(rg_texture{w, h, nullptr, srgb ? 1 : 0, 0});
                } else {
                    L.warnings.push_back("texture not loaded: " + err);
                    tex.push_back(rg_texture{0, 0, nullptr, 0, 0});
                }
            }
            if (tex[k].w) tex_of_obj[o] = (int32_t)k;
        }
        // compact away the textures that failed
        std::vector<rg_texture> used;
        std::vector<int32_t> remap(tex.size(), -1);
        for (size_t k = 0; k < tex.size(); k++)
            if (tex[k].w) {
                remap[k] = (int32_t)used.size();
                used.push_back(tex[k]);
                used.back().rgb = texels[k].data();
            }
        for (int32_t &t : tex_of_obj) t = t >= 0 ? remap[(size_t)t] : -1;
        if (!used.empty())
            rc = rg_scene_set_textures(sc, used.data(), (uint32_t)used.size(), tex_of_obj.data(),
                                       L.obj_tiling.data(), L.uv6.data());
    }
    if (rc) {
        rg_scene_destroy(sc);
        return rc;
    }
    Scene *s = reinterpret_cast<Scene *>(sc);
    s->warnings = L.warnings;
    s->has_camera = L.has_camera;
    *out = sc;
    return RG_OK;
}

extern "C" int rg_scene_warnings(const rg_scene *scene, char *buf, uint64_t cap) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || (!buf && cap)) return set_error(RG_E_INVALID, "null argument");
    std::string all;
    for (const std::string &w : s->warnings) all += w + "\n";
    if (cap) {
        const size_t n = std::min<size_t>(all.size(), (size_t)cap - 1);
        std::memcpy(buf, all.data(), n);
        buf[n] = 0;
    }
    return (int)std::min<size_t>(all.size(), 0x7FFFFFFF);
}

extern "C" int rg_scene_get_camera(const rg_scene *scene, rg_camera *cam) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !cam) return set_error(RG_E_INVALID, "null argument");
    *cam = s->camera;
    return RG_OK;
}

extern "C" int rg_scene_export(const rg_scene *scene, float *v9, float *n9, uint32_t *obj_of_tri, rg_object *objs,
                               uint32_t *n_tris, uint32_t *n_obj) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (n_tris) *n_tris = s->n_tris;
    if (n_obj) *n_obj = (uint32_t)s->objects.size();
    if (v9 && !s->v9.empty()) std::memcpy(v9, s->v9.data(), s->v9.size() * sizeof(float));
    if (n9 && !s->n9.empty()) std::memcpy(n9, s->n9.data(), s->n9.size() * sizeof(float));
    if (obj_of_tri && !s->obj_of_tri.empty())
        std::memcpy(obj_of_tri, s->obj_of_tri.data(), s->obj_of_tri.size() * sizeof(uint32_t));
    if (objs && !s->objects.empty()) std::memcpy(objs, s->objects.data(), s->objects.size() * sizeof(rg_object));
    return RG_OK;
}

extern "C" int rg_scene_export_point_lights(const rg_scene *scene, rg_point_light *lights, uint32_t *n) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !n) return set_error(RG_E_INVALID, "null argument");
    *n = (uint32_t)s->point_lights.size();
    if (lights && !s->point_lights.empty())
        std::memcpy(lights, s->point_lights.data(), s->point_lights.size() * sizeof(rg_point_light));
    return RG_OK;
}

extern "C" int rg_scene_export_media(const rg_scene *scene, rg_medium *media, uint32_t *n_media,
                                     int32_t *global_medium, rg_env *env, float *env_rgb) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (n_media) *n_media = (uint32_t)s->media.size();
    if (global_medium) *global_medium = s->global_medium;
    if (env) {
        env->w = s->env_w;
        env->h = s->env_h;
        env->rgb = nullptr;
        env->intensity = s->env_intensity;
        env->rotation_deg = s->env_rotation_deg;
    }
    if (media && !s->media.empty()) std::memcpy(media, s->media.data(), s->media.size() * sizeof(rg_medium));
    if (env_rgb && !s->env_rgb.empty()) std::memcpy(env_rgb, s->env_rgb.data(), s->env_rgb.size() * sizeof(float));
    return RG_OK;
}
