// tt_kernels.h -- host-side launch wrappers of the gfx950 kernels.  Every
// wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include "tt_device.h"

// ---- scan / compaction (k_misc.hip) ----------------------------------------
// exclusive prefix sum of n uint32 values; tmp must hold scan_tmp_words(n) words
size_t tt_scan_tmp_words(long long n);
hipError_t tt_scan_u32(hipStream_t s, const uint32_t *in, uint32_t *out, long long n, uint32_t *tmp, uint32_t *total);
// ordered compaction of the indices i < n with flags[i] != 0 (u8 flags)
hipError_t tt_compact_u8(hipStream_t s, const uint8_t *flags, long long n, uint32_t *out_idx, uint32_t *count,
                         uint32_t *tmp);
// the same indices in no particular order (flags 16-byte aligned); 2 launches
hipError_t tt_compact_u8_any(hipStream_t s, const uint8_t *flags, long long n, uint32_t *out_idx, uint32_t *count);
hipError_t tt_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, long long n);
hipError_t tt_launch_ledger(hipStream_t s, const float *h, long long n, long long *out);

// ---- initialisation (k_misc.hip) ------------------------------------------
hipError_t tt_launch_init_synthetic_heights(hipStream_t s, const tt_dev &d, uint64_t hseed, double *raw, double *tmp,
                                            const double *wts, float *height);
hipError_t tt_launch_init_reference_heights(hipStream_t s, const tt_dev &d, float *height);
hipError_t tt_launch_init_plate_ids(hipStream_t s, const tt_dev &d, const float *sx, const float *sy, const uint8_t *sid,
                                    int nseeds, uint8_t *ids);
hipError_t tt_launch_asthenosphere(hipStream_t s, const tt_dev &d, const uint8_t *noise, float2 *out);
hipError_t tt_launch_math_eval(hipStream_t s, int fn, const float *in, float *out, long long n);

// ---- stage A/B (k_collide.hip) --------------------------------------------
hipError_t tt_launch_plate_prep(hipStream_t s, const tt_plate_data *pd, tt_plate_step *ps, tt_step_ctl *ctl,
                                long long *fric);
hipError_t tt_launch_advect(hipStream_t s, const tt_dev &d, const uint8_t *ids, const float *height,
                            const tt_plate_step *ps, const tt_step_ctl *ctl_dev, int max_r_host, const uint8_t *typeBitmap,
                            uint32_t *word, uint8_t *cnt, uint32_t *segCount, int *typeCounts, tt_step_ctl *ctl);
hipError_t tt_launch_keys(hipStream_t s, const tt_dev &d, const uint32_t *word, const uint8_t *cnt,
                          const uint32_t *segBase, const uint8_t *ids, const tt_plate_step *ps, const tt_step_ctl *ctl,
                          unsigned long long *keys);
hipError_t tt_launch_type_matrix(hipStream_t s, const tt_dev &d, const int *typeCounts, uint8_t *typeBitmap);
hipError_t tt_launch_resolve(hipStream_t s, const tt_dev &d, const uint32_t *iterp, const uint32_t *word, const uint8_t *ids,
                             const float *height, const tt_plate_step *ps, const uint32_t *divBitmap,
                             const uint8_t *typeBitmap, float *hW, uint8_t *idW, tt_uplift *uplift, uint32_t *accw,
                             uint32_t *hasDiv, long long *fric, uint8_t *srcFlag, uint8_t *contested, int do_uplift);
hipError_t tt_launch_flip(hipStream_t s, uint32_t *divBitmap, const uint32_t *hasDiv);

// ---- stage C (k_uplift.hip) -------------------------------------------------
hipError_t tt_launch_propagate_local(hipStream_t s, const tt_dev &d, const uint32_t *word, const tt_uplift *uplift,
                                     const uint8_t *srcFlag, const uint8_t *contested, tt_prop *prop, int *jfa,
                                     uint8_t *serialFlag);
// cap: sources whose walks k_prop_targets computes (5 cap entries in tcell,
// toff and tslot); slotOf: one word per cell, all 0xFFFFFFFF between calls;
// slotCell: TT_PROP_SLOTS words (k_uplift.hip)
hipError_t tt_launch_propagate_serial(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *count,
                                      const uint32_t *word, const tt_uplift *uplift, const float *height,
                                      const tt_plate_step *ps, tt_prop *prop, int cap, uint32_t *tcell,
                                      uint16_t *toff, uint32_t *tslot, uint32_t *slotOf, uint32_t *slotCell);
int tt_prop_cap(long long cells);
#define TT_PROP_SLOT_WORDS 5120
hipError_t tt_launch_combine_contested(hipStream_t s, const tt_dev &d, const uint32_t *word, const tt_uplift *uplift,
                                       const uint8_t *contested, const uint8_t *srcFlag, tt_prop *prop, int *jfa,
                                       tt_step_ctl *ctl);
hipError_t tt_launch_jfa_pass(hipStream_t s, const tt_dev &d, const int *in, int *out, int step);
hipError_t tt_launch_uplift(hipStream_t s, const tt_dev &d, const int *jfa, const tt_prop *prop, const uint8_t *idW,
                            const tt_plate_step *ps, float *hW, tt_dist *dist);
hipError_t tt_launch_separable_uplift(hipStream_t s, const tt_dev &d, const tt_prop *prop, const uint8_t *idW,
                                      const tt_plate_step *ps, float *hW, tt_dist *dist);

// ---- stage D/E (k_erosion.hip) -----------------------------------------------
int tt_tile_rows_valid(int rows, int H);
// ch / csprev / csout: Kahan compensation arrays of the precision mode (null: plain)
// frozen: per 64x16 tile, unchanged through all thermal passes of the step
// (tt_launch_thermal_frozen); pass: index of the launch within the step;
// two: the launch runs two passes (64x16 tiles, no compensation)
hipError_t tt_launch_thermal(hipStream_t s, const tt_dev &d, int rows, const float *hin, float *hout,
                             float *ch = nullptr, const uint8_t *frozen = nullptr, uint8_t *chg = nullptr,
                             int pass = 0, bool two = false);
// Launches of T thermal passes: passes run in pairs, an odd first pass alone;
// compensated passes (precision mode) one at a time.
inline int tt_thermal_pairs(int T, bool comp) { return comp ? 0 : T / 2; }
inline int tt_thermal_launches(int T, bool comp) { return T - tt_thermal_pairs(T, comp); }
hipError_t tt_launch_thermal_frozen(hipStream_t s, const tt_dev &d, int T, const float *h, float2 *mm,
                                    uint8_t *frozen);
hipError_t tt_launch_hydraulic(hipStream_t s, const tt_dev &d, int rows, int first, int last, int rseed,
                               const float *hin, float *hout, const float *win, float *wout, const float4 *fin,
                               float4 *fout, float2 *vel, const float *sprev, float *sout, float *ch = nullptr,
                               const float *csprev = nullptr, float *csout = nullptr);
hipError_t tt_launch_transport(hipStream_t s, const tt_dev &d, const float *sprev, const float2 *vel, float *sout,
                               const float *csprev = nullptr, float *csout = nullptr);

// ---- stages F/G (k_plates.hip) ----------------------------------------------
hipError_t tt_launch_plate_movement(hipStream_t s, const tt_dev &d, tt_plate_data *pd, const long long *fric,
                                    const uint8_t *noise);
hipError_t tt_launch_accretion_apply(hipStream_t s, const tt_dev &d, const uint32_t *accw, const float *height,
                                     const uint8_t *idsIn, uint8_t *idsOut, uint8_t *accretion);
hipError_t tt_launch_ccl_local(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint32_t *labels, uint8_t *rootFlag,
                               uint32_t *rootCount);
hipError_t tt_launch_ccl_merge(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint32_t *labels);
hipError_t tt_launch_ccl_flatten(hipStream_t s, const uint32_t *locList, const uint32_t *nloc, uint32_t *labels,
                                 uint32_t *rootCount, uint8_t *rootFlag);
hipError_t tt_launch_frag_build(hipStream_t s, const tt_dev &d, const uint32_t *rootList, const uint32_t *nroots,
                                uint32_t nroots_host, const uint32_t *rootCount, const uint8_t *ids, unsigned long long *fragKey,
                                int *maxCount, uint8_t *newId);
hipError_t tt_launch_frag_sort(hipStream_t s, unsigned long long *fragKey, const uint32_t *nroots, int capacity);
// ascending sort of the first *nroots keys (n on the host) for more keys than one block holds
hipError_t tt_sort_keys_large(hipStream_t s, unsigned long long *key, const uint32_t *nroots, uint32_t n);
hipError_t tt_launch_frag_assign(hipStream_t s, const tt_dev &d, const unsigned long long *fragKey,
                                 const uint32_t *nroots, const int *maxCount, const uint8_t *ids, tt_plate_data *pd,
                                 uint8_t *newId);
hipError_t tt_launch_relabel(hipStream_t s, const tt_dev &d, uint32_t *labels, const uint8_t *newId,
                             uint8_t *ids, uint8_t *unFlag);
hipError_t tt_launch_fill_vote(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                               const uint8_t *ids, uint8_t *vote);
hipError_t tt_launch_fill_apply(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                                const uint8_t *vote, uint8_t *ids, uint32_t *nextList, uint32_t *nextN);
hipError_t tt_launch_neighbor_matrix(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint8_t *nbm);
hipError_t tt_launch_merge_decide(hipStream_t s, const tt_dev &d, const uint8_t *nbm, const tt_plate_data *pd,
                                  int *mergeIds, uint8_t *typeBitmap, tt_step_ctl *ctl);
hipError_t tt_launch_count(hipStream_t s, const tt_dev &d, const int *mergeIds, const uint8_t *idsIn, uint8_t *idsOut,
                           const float *height, const uint8_t *noise, long long *acc, int split_mode);
hipError_t tt_launch_finalize(hipStream_t s, const tt_dev &d, tt_plate_data *pd, const long long *acc, int mode,
                              tt_step_ctl *ctl);
// split
hipError_t tt_launch_split_prepare(hipStream_t s, const tt_dev &d, const uint8_t *ids, int plate, const float *height,
                                   float *eff, uint8_t *flag, tt_step_ctl *ctl);
int tt_split_tile_h(const tt_dev &d);
hipError_t tt_launch_split_tiles(hipStream_t s, const tt_dev &d, const uint8_t *ids, int plate, uint8_t *flag);
hipError_t tt_launch_relax_tiles(hipStream_t s, const tt_dev &d, const uint32_t *tiles, unsigned ntiles,
                                 const uint8_t *ids, int plate, const float *height, float *eff, uint32_t *rchg,
                                 uint8_t *act, int round0, int rounds);
hipError_t tt_launch_backtrack(hipStream_t s, const tt_dev &d, const float *eff, uint8_t *ids, int px, int py,
                               int newId, tt_step_ctl *ctl);
hipError_t tt_launch_flood_cc(hipStream_t s, const tt_dev &d, uint8_t *ids, int oldId, int newId, uint32_t *lab,
                              uint8_t *rootFlag, uint32_t *list, uint32_t *nlist, uint32_t *scanTmp,
                              const tt_step_ctl *ctl);
hipError_t tt_launch_split_finalize(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                                    uint8_t *ids, int oldId, int newId, tt_plate_data *pd);
hipError_t tt_launch_probe(hipStream_t s, int which, const float4 *a, float4 *b, float *sink, long long n);
// counts (into *bad, host memory) the numerators for which tt_div_const(x, L, rL) differs from x / L
hipError_t tt_check_divisor(hipStream_t s, float L, float rL, uint32_t *bad);
// counts (into bad[0], host memory) the arguments in [2^-96, FLT_MAX] where tt_sqrt_normal differs from
// sqrtf, and (into bad[1]) the floats where tt_f2u32_hw differs from tt_f2u32_sat
hipError_t tt_check_sqrt(hipStream_t s, uint32_t bad[2]);
