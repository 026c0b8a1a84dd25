/* A plain C client of the drop-in's C ABI (include/beamlab_gpu.h), as the reference's host would call it: the C
 * compiler sees only the header, the linker only libbeamlab_gpu.so. Built and run by tests/test_abi_client.py.
 *
 *   client info                 the library's version, default settings and body options, error strings (no GPU)
 *   client prep IN OUT          bl_finalize + bl_stabilize of a body read from IN, written to OUT (no GPU)
 *   client step IN OUT N        a world with the ground box and the body of IN, N substeps, the nodes and beams to OUT
 *
 * A body file: uint32 n_nodes, n_beams, then n_nodes bl_node, n_nodes bl_node_info, n_beams bl_beam. */
#include "beamlab_gpu.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t n, nb;
    bl_node* nodes;
    bl_node_info* info;
    bl_beam* beams;
} body_t;

static int read_body(const char* path, body_t* b) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int ok = fread(&b->n, 4, 1, f) == 1 && fread(&b->nb, 4, 1, f) == 1;
    if (ok) {
        b->nodes = malloc(sizeof(bl_node) * (b->n ? b->n : 1));
        b->info = malloc(sizeof(bl_node_info) * (b->n ? b->n : 1));
        b->beams = malloc(sizeof(bl_beam) * (b->nb ? b->nb : 1));
        ok = fread(b->nodes, sizeof(bl_node), b->n, f) == b->n && fread(b->info, sizeof(bl_node_info), b->n, f) == b->n &&
             fread(b->beams, sizeof(bl_beam), b->nb, f) == b->nb;
    }
    fclose(f);
    return ok ? 0 : -1;
}

static int write_body(const char* path, const body_t* b) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int ok = fwrite(&b->n, 4, 1, f) == 1 && fwrite(&b->nb, 4, 1, f) == 1 &&
             fwrite(b->nodes, sizeof(bl_node), b->n, f) == b->n && fwrite(b->info, sizeof(bl_node_info), b->n, f) == b->n &&
             fwrite(b->beams, sizeof(bl_beam), b->nb, f) == b->nb;
    fclose(f);
    return ok ? 0 : -1;
}

static int fail(const char* what, int rc) {
    fprintf(stderr, "%s: %d (%s)\n", what, rc, bl_error_string(rc));
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (!strcmp(argv[1], "info")) {
        bl_settings s;
        bl_body_opts o;
        bl_default_settings(&s);
        bl_default_body_opts(&o);
        printf("abi %d\n", bl_abi_version());
        printf("dt %g collision_hz %g theta %g dissipation %g kill_y %g\n", s.dt, s.collision_hz, s.frame_theta,
               s.frame_dissipation, s.kill_y);
        printf("collision_radius %g hull_depth %g face_skin %g\n", o.collision_radius, o.hull_depth, o.face_skin);
        printf("error %s\n", bl_error_string(BL_ERR_UNSUPPORTED));
        return 0;
    }
    if (argc < 4) return 2;
    body_t b;
    if (read_body(argv[2], &b)) return fail("read", -1);
    if (!strcmp(argv[1], "prep")) {
        int rc = bl_finalize(b.nodes, b.n, b.beams, b.nb, NULL, 0);
        if (rc) return fail("bl_finalize", rc);
        int32_t changed = 0;
        rc = bl_stabilize(b.nodes, b.n, b.beams, b.nb, 0.0005f, 0.6f, 0.5f, &changed);
        if (rc) return fail("bl_stabilize", rc);
        printf("stabilized %d\n", changed);
        return write_body(argv[3], &b) ? fail("write", -1) : 0;
    }
    if (!strcmp(argv[1], "step") && argc >= 5) {
        bl_settings s;
        bl_default_settings(&s);
        bl_ctx* ctx = NULL;
        int rc = bl_create(&ctx, &s);
        if (rc) return fail("bl_create", rc);
        const float c[3] = {0.0f, -0.5f, 0.0f}, half[3] = {100.0f, 0.5f, 100.0f}, q[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        bl_box ground;
        if ((rc = bl_box_from_quat(c, half, q, BL_SURF_CONCRETE, &ground))) return fail("bl_box_from_quat", rc);
        if ((rc = bl_set_statics(ctx, &ground, 1))) return fail("bl_set_statics", rc);
        bl_body_desc d;
        memset(&d, 0, sizeof(d));
        d.nodes = b.nodes, d.info = b.info, d.n_nodes = b.n, d.beams = b.beams, d.n_beams = b.nb;
        bl_default_body_opts(&d.opts);
        uint32_t id = 0;
        if ((rc = bl_add_body(ctx, &d, &id))) return fail("bl_add_body", rc);
        if ((rc = bl_step(ctx, atoi(argv[4])))) return fail("bl_step", rc);
        if ((rc = bl_read_body(ctx, id, b.nodes, b.beams))) return fail("bl_read_body", rc);
        bl_stats st;
        if ((rc = bl_get_stats(ctx, &st))) return fail("bl_get_stats", rc);
        printf("substeps %lld static_contacts %lld broken %lld\n", (long long)st.substeps, (long long)st.static_contacts,
               (long long)st.broken_beams);
        bl_destroy(ctx);
        return write_body(argv[3], &b) ? fail("write", -1) : 0;
    }
    return 2;
}
