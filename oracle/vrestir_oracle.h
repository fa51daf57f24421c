/* vrestir_oracle.h -- TEST INFRASTRUCTURE ONLY: C interface of the CPU oracle.
 *
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * only, always as the checker. The product library never links it. */
#ifndef VRESTIR_ORACLE_H
#define VRESTIR_ORACLE_H
#include "../include/vrestir.h"

#ifdef __cplusplus
extern "C" {
#endif
void vrestir_default_params_oracle(vrestir_params* p);
void* orc_create(void);
void orc_destroy(void* h);
int orc_set_params(void* h, const vrestir_params* p);
int orc_set_volume(void* h, const vrestir_volume_desc* d);
int orc_set_lights(void* h, const vrestir_point_light* lights, int n, const float env[3]);
int orc_set_surfaces(void* h, const vrestir_surfaces_desc* desc);
void orc_surface_hit(void* h, int n, const float* rays, float* out);
void orc_surface_shade(void* h, int variant, int n, const float* rays, const uint32_t* seeds, float* out);
void orc_sample_lights(void* h, int kinds, int n, const float* points, const uint32_t* seeds, float* out);
void orc_emissive_eval(void* h, int n, const float* points, const int* tris, const float* uvs, float* out);
int orc_set_camera(void* h, const vrestir_camera* c, int width, int height);
int orc_add_volume_frame(void* h, const vrestir_volume_desc* d);
int orc_set_volume_frame(void* h, int index, int animating);
int orc_set_freeze_frame(void* h, int freeze);
int orc_set_temperature(void* h, const float* t, int nx, int ny, int nz, const float* indexToWorld, int fp16,
                        float LeScale, float cutoff, float scale);
int orc_set_blackbody_lut(void* h, const float* rgb, int rows);
int orc_set_slot(void* h, int slot, const float* v, int nx, int ny, int nz, const float* indexToWorld, int fp16);
int orc_set_env_map(void* h, int w, int hgt, const float* rgb, float intensity, const float tint[3],
                    const float rotationDegrees[3], int impDim, int impSpp);
int orc_set_env_rotation(void* h, const float deg[3]);
int orc_set_env_intensity(void* h, float intensity, const float tint[3]);
void orc_env_rotation(void* h, int prev, float toWorld[9], float toLocal[9]);
void orc_env_eval(void* h, const float dir[3], float out[3]);
void orc_env_sample(void* h, float u, float v, float out[8]);
float orc_env_pdf(void* h, const float dir[3]);
int orc_env_importance(void* h, int mip, float* out);
void orc_latlong(const float dir[3], float uv[2]);
void orc_dir_to_oct(const float dir[3], float uv[2]);
void orc_emission(void* h, const float p[3], float out[4]);
void orc_blackbody(void* h, float temperature, float out[3]);
int orc_execute(void* h, int width, int height, float* outRGBA, int rowBegin, int rowEnd, int nthreads);
int orc_prepare_frame(void* h, int width, int height, int nthreads);
/* Restrict the following stages to image rows [rowBegin, rowEnd) (after prepare). */
int orc_set_rows(void* h, int rowBegin, int rowEnd);
/* Oracle-only buffer id: the motion vectors TemporalReuse writes (float2 per pixel). */
#define ORC_BUF_MVEC 100
int orc_run_stage(void* h, int stage, int round);
int orc_end_frame(void* h);
int orc_read_buffer(void* h, int id, void* dst, size_t bytes);
int orc_write_buffer(void* h, int id, const void* src, size_t bytes);
int orc_frame_count(void* h);
int orc_total_round_id(void* h);
void orc_stage_seconds(void* h, double out[6]);
int orc_track_batch(void* h, int mode, int mip, int linear, float tStepScale, int numSamples, int n,
                    const float* rays, const uint32_t* seeds, float* out);
void orc_rng_stream(uint32_t px, uint32_t py, uint32_t sampleNumber, int n, uint32_t* out);
uint32_t orc_morton(uint32_t x, uint32_t y);
void orc_splitmix64(uint64_t seed, int n, uint64_t* out);
void orc_xoshiro128ss(const uint32_t seed[4], int n, uint32_t* out);
float orc_round_to_half(float v);
void orc_neighbor_offsets(int sampleCount, float radius, int frameSeed, int sampler, int* out);
int orc_grid_info(void* h, int slot, int* dims, float* maxValue, int* numBricks);
int orc_brick_bounds(void* h, int slot, int x, int y, int z, float* out4);
float orc_voxel(void* h, int slot, int x, int y, int z);
int orc_final_costs(void* h, long long* out);
float orc_sample_phase(float g, const float wo[3], float u0, float u1, float wi[3]);
float orc_phase_hg(float cosTheta, float g);
void orc_oct_to_dir(float u, float v, float out[3]);
void orc_light_uv_encode(const float dir[3], float uv[2], int* id);
void orc_light_uv_decode(const float uv[2], int id, float out[3]);
void orc_wi_dist_encode(const float wiDist[4], float out[3]);
void orc_wi_dist_decode(const float in[3], float out[4]);
int orc_pack_sampled_pixel(int pixel, int tag, int maxIndirect);
void orc_unpack_sampled_pixel(int s, int* tag, int* maxIndirect);
int orc_wrs(int n, const float* w, const float* M, float cap, uint32_t px, uint32_t py, uint32_t sn, float* runningSum,
            float* outM);
int orc_camera_ray(void* h, int px, int py, int W, int H, float out[3]);
int orc_reproject(void* h, const float p[3], int W, int H, float out[2]);
int orc_trace_bricks(void* h, int slot, const float ray[8], int vertexCenter, int maxOut, float* out);
/* learned proposal */
int orc_set_learned(void* h, int enable, int train, int bx, int by, int steps, float lr, float alpha, int judge);
int orc_learned_set_weights(void* h, const float* w, int n);
int orc_learned_get_weights(void* h, float* w, int n);
void orc_learned_stats(void* h, float out[5]);
int orc_learned_raw(void* h, int px, int py, int W, int H, float* raw);
void orc_mlp_encode(const float* raw, float* out);
void orc_mlp_forward(const float* prm, const float* raw, int n, float alpha, int emulate, float* z, uint16_t* cdf);
double orc_mlp_loss(const float* prm, const float* recs, int n, float alpha, int emulate);
void orc_mlp_backward(const float* prm, const float* recs, int n, float alpha, float scale, int emulate, double* grad,
                      double* stats);
void orc_mlp_adam(float* prm, float* m, float* v, const float* g, int n, int t, float lr);
float orc_mlp_sample(const uint16_t* cdf, float u, int* bin, float* g);
float orc_mlp_mass_above(const uint16_t* cdf, float x);
int orc_mlp_train_slot(int x, int y, int W, int H, int bx, int by, int frame);
#ifdef __cplusplus
}
#endif
#endif
