/* Oracle (test infrastructure only): CPU pointing expansion for the synthetic ring scan. */
#ifndef ORACLE_SCAN_H
#define ORACLE_SCAN_H
#include <stdint.h>

typedef struct {
    double fsample;        /* Hz */
    double spin_period;    /* s per spin revolution */
    double prec_period;    /* s per full turn of the spin-axis longitude */
    double opening_turns;  /* boresight opening angle, turns (85 deg = 85/360) */
    double lon0_turns;     /* spin-axis longitude at sample 0, turns */
} oscan_params;

typedef struct {
    double psi_turns;          /* polarisation angle from the scan direction */
    double open_offset_turns;  /* cross-scan offset added to the opening angle */
    double spin_offset_turns;  /* in-scan offset added to the spin phase */
} oscan_detector;

void oscan_sincos_turns(double u, double *s, double *c);
double oscan_atan2_pos(double y, double x);
int64_t oscan_vec2pix_nest(int64_t nside, double x, double y, double z);
void oscan_expand(const oscan_params *sp, const oscan_detector *det, int64_t nside,
                  int64_t i0, int64_t n, int32_t *pix, float *w);

#endif
