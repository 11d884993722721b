/*
 * fl_oracle.h -- CPU restatement of the search-side RFI flagger
 * (gpu_flagger.cu:113-790, loop body :1294-1400) on a
 * [nbeam][ntime][nchan] uint8 block.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Shapes: ntime a multiple of 32 * 128 (ppspec bins), nchan a multiple of 32
 * and above 256 + 21 for the channel-mask functions.
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_fl_spectrum(const uint8_t *data, int nb, int nt, int nc, float *spec);
void orc_fl_varspec(const uint8_t *data, int nb, int nt, int nc, const float *spec, float *var);
void orc_fl_varspec_exact(const uint8_t *data, int nb, int nt, int nc, const float *spec, float *var);
void orc_fl_maxspec(const uint8_t *data, int nb, int nt, int nc, float *mx);
void orc_fl_ppspec(const uint8_t *data, int nb, int nt, int nc, float *pp);
void orc_fl_scaley(uint8_t *data, int nb, int nt, int nc, const float *spec, const float *var);
void orc_fl_sumpulse(uint8_t *data, long n, const float *summand);
void orc_fl_make_ts(const uint8_t *data, int nb, int nt, int nc, float *ts);
void orc_fl_flag(uint8_t *data, int nb, int nt, int nc, const int *midx, int n_mask, const float *bpwr);
void orc_fl_flagts(uint8_t *data, int nb, int nt, int nc, const int *midx, int n_mask, const float *bpwr);
uint64_t orc_fl_mix(uint64_t z);
void orc_fl_genrand(uint8_t *repval, long n, uint32_t seed);
void orc_fl_fix_zspec(float *s0, float *v0, int nbc, int naver);

float orc_fl_medval(const float *a, int n);
void orc_fl_simple_channflag(const float *spec, int nb, int nc, float thr, int *mask);
void orc_fl_simple_tsflag(const float *ts, int nb, int nt, float thr, int *mask);
void orc_fl_channflag(const float *spec, int nb, int nc, float thr, const int *perm, int nperm,
                      int *mask);
int orc_fl_gather(const int *mask, long n, int *idx);
void orc_fl_calc_bpwr(const float *spec, int nb, int nc, float *bpwr);
void orc_fl_median_calc(float *arr, int nb, int nc);
float orc_fl_step(uint8_t *data, int nb, int nt, int nc, const float *spec0, const float *var0,
                  const uint8_t *repval, float prev_tpwr, float mod_thresh, float thresh,
                  float tsthresh, int pwr, int mkrand, int tsflag, const int *perm, int nperm,
                  int exact_var, float *bpwr, int *n_mask, int *n_tsmask);

#ifdef __cplusplus
}
#endif
#endif
