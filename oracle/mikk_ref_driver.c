/* Test infrastructure: drives the reference's vendored mikktspace
 * (External/mikktspace/mikktspace.c, compiled where it lies into
 * oracle/_ref/libmikk_ref.so by oracle/build.py) over an indexed triangle mesh
 * the way the reference's import does (TangentSpace.cpp:14-89): faces of
 * three corners from an index list, attributes read per vertex, and
 * genTangSpaceDefault's (tangent, sign) written into the shared vertex, so the
 * last face in index order sets it. */
#include <stdint.h>
#include <string.h>

#include "mikktspace.h"

typedef struct {
    const float *pos, *nrm, *uv;
    const uint32_t* idx;
    int faces;
    float* tan;
} mesh_t;

static int n_faces(const SMikkTSpaceContext* c) { return ((const mesh_t*)c->m_pUserData)->faces; }
static int n_verts(const SMikkTSpaceContext* c, const int f) {
    (void)c;
    (void)f;
    return 3;
}
static void get3(const float* a, uint32_t v, float out[]) {
    out[0] = a[3 * (size_t)v];
    out[1] = a[3 * (size_t)v + 1];
    out[2] = a[3 * (size_t)v + 2];
}
static void get_pos(const SMikkTSpaceContext* c, float out[], const int f, const int k) {
    const mesh_t* m = (const mesh_t*)c->m_pUserData;
    get3(m->pos, m->idx[3 * f + k], out);
}
static void get_nrm(const SMikkTSpaceContext* c, float out[], const int f, const int k) {
    const mesh_t* m = (const mesh_t*)c->m_pUserData;
    get3(m->nrm, m->idx[3 * f + k], out);
}
static void get_uv(const SMikkTSpaceContext* c, float out[], const int f, const int k) {
    const mesh_t* m = (const mesh_t*)c->m_pUserData;
    const uint32_t v = m->idx[3 * f + k];
    out[0] = m->uv[2 * (size_t)v];
    out[1] = m->uv[2 * (size_t)v + 1];
}
static void set_basic(const SMikkTSpaceContext* c, const float t[], const float sign, const int f, const int k) {
    mesh_t* m = (mesh_t*)c->m_pUserData;
    float* o = m->tan + 4 * (size_t)m->idx[3 * f + k];
    o[0] = t[0];
    o[1] = t[1];
    o[2] = t[2];
    o[3] = sign;
}

/* Returns genTangSpaceDefault's result (1 = ok); tangents start at zero. */
int mikk_ref_tangents(const float* pos, const float* nrm, const float* uv, uint32_t vertex_count,
                      const uint32_t* idx, uint32_t index_count, float* tangents) {
    memset(tangents, 0, sizeof(float) * 4 * (size_t)vertex_count);
    mesh_t m = {pos, nrm, uv, idx, (int)(index_count / 3), tangents};
    SMikkTSpaceInterface iface;
    memset(&iface, 0, sizeof(iface));
    iface.m_getNumFaces = n_faces;
    iface.m_getNumVerticesOfFace = n_verts;
    iface.m_getPosition = get_pos;
    iface.m_getNormal = get_nrm;
    iface.m_getTexCoord = get_uv;
    iface.m_setTSpaceBasic = set_basic;
    SMikkTSpaceContext ctx;
    memset(&ctx, 0, sizeof(ctx));
    ctx.m_pInterface = &iface;
    ctx.m_pUserData = &m;
    return genTangSpaceDefault(&ctx);
}
