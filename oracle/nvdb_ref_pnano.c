/*
 * Checker-side reader: the reference's PNanoVDB accessor (Data/Shaders/VPT/PNanoVDB.glsl, C branch) compiled
 * from the reference tree by oracle/Makefile. TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/nvdb_ref, which
 * writes golden vectors for the tests.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#define PNANOVDB_C
#define PNANOVDB_ADDRESS_32
#include "PNanoVDB.glsl"

/* pnanovdb_readaccessor_get_value_address + pnanovdb_read_float at integer coordinates (VptUtils.glsl:515-523). */
void ref_lookup_ijk(const uint32_t* words, int n, const int32_t* ijk, float* out) {
    pnanovdb_buf_t buf = pnanovdb_make_buf((uint32_t*)words, 0);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_root_handle_t root = pnanovdb_tree_get_root(buf, pnanovdb_grid_get_tree(buf, grid));
    pnanovdb_readaccessor_t acc;
    pnanovdb_readaccessor_init(&acc, root);
    for (int i = 0; i < n; i++) {
        pnanovdb_coord_t c;
        c.x = ijk[3 * i];
        c.y = ijk[3 * i + 1];
        c.z = ijk[3 * i + 2];
        pnanovdb_address_t a = pnanovdb_readaccessor_get_value_address(PNANOVDB_GRID_TYPE_FLOAT, buf, &acc, &c);
        out[i] = pnanovdb_read_float(buf, a);
    }
}

/* Nearest sampleCloudRaw at world (NanoVDB grid) coordinates: world_to_indexf, floor, lookup. */
void ref_lookup_world(const uint32_t* words, int n, const float* xyz, float* out, int32_t* ijk_out) {
    pnanovdb_buf_t buf = pnanovdb_make_buf((uint32_t*)words, 0);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_root_handle_t root = pnanovdb_tree_get_root(buf, pnanovdb_grid_get_tree(buf, grid));
    pnanovdb_readaccessor_t acc;
    pnanovdb_readaccessor_init(&acc, root);
    for (int i = 0; i < n; i++) {
        pnanovdb_vec3_t p;
        p.x = xyz[3 * i];
        p.y = xyz[3 * i + 1];
        p.z = xyz[3 * i + 2];
        pnanovdb_vec3_t idx = pnanovdb_grid_world_to_indexf(buf, grid, &p);
        pnanovdb_coord_t c;
        c.x = (int32_t)floorf(idx.x);
        c.y = (int32_t)floorf(idx.y);
        c.z = (int32_t)floorf(idx.z);
        ijk_out[3 * i] = c.x;
        ijk_out[3 * i + 1] = c.y;
        ijk_out[3 * i + 2] = c.z;
        pnanovdb_address_t a = pnanovdb_readaccessor_get_value_address(PNANOVDB_GRID_TYPE_FLOAT, buf, &acc, &c);
        out[i] = pnanovdb_read_float(buf, a);
    }
}

/* Fp16 grids: the same lookups through pnanovdb_root_fp16_read_float (code * quantum + minimum at leaf level). */
void ref_lookup_ijk16(const uint32_t* words, int n, const int32_t* ijk, float* out) {
    pnanovdb_buf_t buf = pnanovdb_make_buf((uint32_t*)words, 0);
    pnanovdb_grid_handle_t grid = {pnanovdb_address_null()};
    pnanovdb_root_handle_t root = pnanovdb_tree_get_root(buf, pnanovdb_grid_get_tree(buf, grid));
    pnanovdb_readaccessor_t acc;
    pnanovdb_readaccessor_init(&acc, root);
    for (int i = 0; i < n; i++) {
        pnanovdb_coord_t c;
        c.x = ijk[3 * i];
        c.y = ijk[3 * i + 1];
        c.z = ijk[3 * i + 2];
        pnanovdb_uint32_t level = 0u;
        pnanovdb_address_t a =
            pnanovdb_readaccessor_get_value_address_and_level(PNANOVDB_GRID_TYPE_FP16, buf, &acc, &c, &level);
        out[i] = pnanovdb_root_fp16_read_float(buf, a, &c, level);
    }
}
