/*
 * orc_frame.c -- oracle: 3-D corotational frame element, lumped masses and the
 * nonlinear Newmark driver.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 */
#include "cb_oracle.h"

#include <stdio.h>

void orc_stiff_fr(const orc_dims *d, double *ss, const orc_model *m, const double *llength,
                  const double *defllen_ip, const double *c1_ip, const double *c2_ip,
                  const double *c3_ip, const double *ef_ip, const double *efFE_ip)
{
    (void)d; (void)ss; (void)m; (void)llength; (void)defllen_ip; (void)c1_ip; (void)c2_ip;
    (void)c3_ip; (void)ef_ip; (void)efFE_ip;
}

void orc_forces_fr(const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                   const double *efFE_ip, double *efFE_i, const double *dd,
                   const orc_model *m, const double *llength, const double *defllen_ip,
                   const double *c1_ip, const double *c2_ip, const double *c3_ip,
                   const double *c1_i, const double *c2_i, const double *c3_i)
{
    (void)d; (void)f; (void)ef_ip; (void)ef_i; (void)efFE_ip; (void)efFE_i; (void)dd; (void)m;
    (void)llength; (void)defllen_ip; (void)c1_ip; (void)c2_ip; (void)c3_ip; (void)c1_i;
    (void)c2_i; (void)c3_i;
}

void orc_mass(const orc_dims *d, double *sm, const orc_model *m, const double *x,
              double *llength_out)
{
    (void)d; (void)sm; (void)m; (void)x; (void)llength_out;
}

