/*
 * hii_tmpl.h -- the find_HII_bubbles restatement, instantiated by hii.c for the faithful
 * FP32 variant (REAL = float) and the FP64 oracle (REAL = double).
 * TEST INFRASTRUCTURE ONLY.
 *
 * Follows src/core/find_HII_bubbles.c:86-710 line by line for the configurations the
 * library supports (no mini-haloes).  With Flag_IncludeRecombinations the post-loop lookup uses
 * the tables set by oracle_rec_set_tables.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* Returns the number of radii processed (all of them unless max_radii > 0). */
static int CAT(hii_run_, SUFFIX)(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                 double* margin, int max_radii, double* seconds, double* seconds_fwd)
{
  const int N = p->ReionGridDim;
  const long H = N / 2 + 1;
  const long Nc = (long)N * N * H;
  const long n_real = (long)N * N * N;

  const int recomb = p->Flag_IncludeRecombinations;
  if (recomb && (!g->N_rec || !g->Gamma12 || !g->residual_xH || !g->clumping_factor || !g->t_resp))
    return MXR_ERR_INVALID_ARG;
  if (p->ReionFilterType < 0 || p->ReionFilterType > 2)
    return MXR_ERR_FILTER;

  /* find_HII_bubbles.c:90-140 */
  const double box_size = p->BoxSize;
  const double pixel_volume = pow(box_size / (double)N, 3);
  double cell_length_factor = L_FACTOR;
  const double total_n_cells = pow((double)N, 3);
  const int flag_ReionUVBFlag = p->ReionUVBFlag;
  const double ReionEfficiency = p->ReionEfficiency;
  const double ReionNionPhotPerBary = p->ReionNionPhotPerBary;
  const double UnitLength_in_cm = p->UnitLength_in_cm;
  const double UnitMass_in_g = p->UnitMass_in_g;
  const double UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  const double sfr_timescale = p->ReionSfrTimescale * oracle_hubble_time(p);
  const double redshift = p->redshift;
  double TKd = p->TK_recfast;
  if (TKd < 0)
    TKd = oracle_T_RECFAST(p->TablesForXHeatingDir, (float)redshift);
  if (TKd < 0)
    return MXR_ERR_TABLES;
  const float TK = (float)TKd;
  const float cT_ad = cT_approx((float)redshift);
  if ((box_size / (double)N) < 1.0)
    cell_length_factor = 1.0;

  const int spin = p->Flag_IncludeSpinTemp;
  const int bh = p->Flag_BHFeedback;
  /* grids: deltax, stars, weighted_sfr, [x_e], [effective_bhm, effective_bhar], [N_rec] */
  const int ngrid = 3 + (spin ? 1 : 0) + (bh ? 2 : 0) + (recomb ? 1 : 0);
  const int IXE = 3, IBHM = 3 + (spin ? 1 : 0), IBHAR = 4 + (spin ? 1 : 0), INREC = ngrid - 1;
  const float* src[7] = {g->deltax, g->stars, g->weighted_sfr, NULL, NULL, NULL, NULL};
  if (spin)
    src[IXE] = g->x_e_box;
  if (bh) {
    src[IBHM] = g->effective_bhm;
    src[IBHAR] = g->effective_bhar;
  }
  if (recomb)
    src[INREC] = g->N_rec;
  float* Gamma12 = g->Gamma12;
  float* N_rec = g->N_rec;
  /* :128-130 */
  const float zstep = (float)(p->prev_redshift - redshift);
  const float fabs_dtdz = (float)fabs(oracle_dtdz(p->OmegaM, p->OmegaLambda, p->Hubble_h, (float)redshift) /
                                      p->Hubble_h);
  float* xH = g->xH;
  float* r_bubble = g->r_bubble;
  float* z_in = g->z_at_ionization;
  float* J_21_at_ionization = g->J_21_at_ionization;
  float* temp_kinetic_all_gas = g->temp_kinetic_all_gas;
  const float* Tk_box = g->Tk_box;
  const float* deltax = g->deltax;
  const float* weighted_sfr = g->weighted_sfr;

  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);

  /* :150-158 */
  for (long ii = 0; ii < n_real; ii++)
    xH[ii] = 1.0;
  for (long ii = 0; ii < n_real; ii++)
    r_bubble[ii] = 0.0;
  if (margin)
    for (long ii = 0; ii < n_real; ii++)
      margin[ii] = -HUGE_VAL;

  /* :163-229 forward transforms (out of place, padded real input) */
  REAL* unfiltered[7] = {0};
  REAL* filtered[7] = {0};
  REAL* rin = (REAL*)malloc(sizeof(REAL) * 2 * Nc);
  for (int gi = 0; gi < ngrid; gi++) {
    unfiltered[gi] = (REAL*)malloc(sizeof(REAL) * 2 * Nc);
    filtered[gi] = (REAL*)malloc(sizeof(REAL) * 2 * Nc);
    for (long ii = 0; ii < 2 * Nc; ii++)
      rin[ii] = (REAL)src[gi][ii];
    CAT(r2c3d_, SUFFIX)(N, rin, (CAT(cplx_, SUFFIX)*)unfiltered[gi]);
    /* :235-253  spectrum /= total_n_cells (a double divisor) */
    for (long ii = 0; ii < 2 * Nc; ii++)
      unfiltered[gi][ii] = (REAL)((double)unfiltered[gi][ii] / total_n_cells);
  }
  free(rin);
  if (seconds_fwd) {
    struct timespec tf;
    clock_gettime(CLOCK_MONOTONIC, &tf);
    *seconds_fwd = (double)(tf.tv_sec - t0.tv_sec) + 1e-9 * (double)(tf.tv_nsec - t0.tv_nsec);
  }

  /* :256-271 */
  double ReionRBubbleMax;
  if (p->Flag_EvolvingReionRBubbleMax) {
    if (redshift > 6.)
      ReionRBubbleMax = 25.483241248322766;
    else
      ReionRBubbleMax = 112. * pow((1. + redshift) / 5., -4.4);
  } else {
    ReionRBubbleMax = p->ReionRBubbleMax;
  }
  const double ReionRBubbleMin = p->ReionRBubbleMin;
  double R = fmin(ReionRBubbleMax, L_FACTOR * box_size);
  const double ReionDeltaRFactor = p->ReionDeltaRFactor;
  const double ReionGammaHaloBias = p->ReionGammaHaloBias;

  int flag_last_filter_step = 0;
  int n_radii = 0;

  while (!flag_last_filter_step) {
    /* :281-285 */
    if (((R / ReionDeltaRFactor) <= (cell_length_factor * box_size / (double)N)) ||
        ((R / ReionDeltaRFactor) <= ReionRBubbleMin)) {
      flag_last_filter_step = 1;
      R = cell_length_factor * box_size / (double)N;
    }
    if (max_radii > 0 && n_radii >= max_radii)
      break;
    n_radii++;
    oracle_progress("radius", n_radii, 0);

    /* :292-348 copy, then filter unless this is the last step */
    for (int gi = 0; gi < ngrid; gi++) {
      memcpy(filtered[gi], unfiltered[gi], sizeof(REAL) * 2 * Nc);
      if (!flag_last_filter_step)
        CAT(filter_, SUFFIX)(filtered[gi], 0, N, N, box_size, (FILTER_R_T)R, p->ReionFilterType);
      /* :351-370 in-place c2r */
      CAT(c2r3d_, SUFFIX)(N, (CAT(cplx_, SUFFIX)*)filtered[gi]);
    }

    /* :372-422 aliasing clamp */
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < N; ix++)
      for (int iy = 0; iy < N; iy++)
        for (int iz = 0; iz < N; iz++) {
          const long ip = iz + 2 * H * (iy + (long)N * ix);
          filtered[0][ip] = FMAXR(filtered[0][ip], (REAL)(-1 + REL_TOL));
          for (int gi = 1; gi < 3; gi++) {
            filtered[gi][ip] = FMAXR(filtered[gi][ip], (REAL)0.0);
            if (filtered[gi][ip] < (REAL)ABS_TOL)
              filtered[gi][ip] = 0;
          }
          if (spin) {
            filtered[IXE][ip] = FMAXR(filtered[IXE][ip], (REAL)0.0);
            if (filtered[IXE][ip] < (REAL)ABS_TOL)
              filtered[IXE][ip] = 0;
            filtered[IXE][ip] = FMINR(filtered[IXE][ip], (REAL)0.999);
          }
          if (bh)
            for (int gi = IBHM; gi <= IBHAR; gi++) {
              filtered[gi][ip] = FMAXR(filtered[gi][ip], (REAL)0.0);
              if (filtered[gi][ip] < (REAL)ABS_TOL)
                filtered[gi][ip] = 0;
            }
          if (recomb) {
            filtered[INREC][ip] = FMAXR(filtered[INREC][ip], (REAL)0.0);
            if (filtered[INREC][ip] < (REAL)ABS_TOL)
              filtered[INREC][ip] = 0;
          }
        }

    /* :427-442 per-radius prefactors */
    const double M_mean = oracle_RtoM(p, R);
    const double R_cubed = R * R * R;
    double Gamma_R_prefactor =
      (1.0 + redshift) * (1.0 + redshift) * R * UnitLength_in_cm / pixel_volume * pow(UnitLength_in_cm, -3.);
    Gamma_R_prefactor *= (UnitMass_in_g / UnitTime_in_s) / PROTONMASS * ReionNionPhotPerBary;
    double Gamma_R_prefactor_BH = Gamma_R_prefactor;
    const float J_21_aux_constant = (float)(Gamma_R_prefactor * PLANCK * 1e21 / (4.0 * M_PI) * ReionGammaHaloBias /
                                            sfr_timescale * p->ReionAlphaUV);
    const float J_21_aux_constant_BH = (float)(Gamma_R_prefactor_BH * PLANCK * 1e21 / (4.0 * M_PI) *
                                               ReionGammaHaloBias / sfr_timescale * p->ReionAlphaUVBH);
    Gamma_R_prefactor *= SIGMA_HI * p->ReionAlphaUV / (p->ReionAlphaUV + 2.75) * 1e12;
    Gamma_R_prefactor_BH *= SIGMA_HI * p->ReionAlphaUVBH / (p->ReionAlphaUVBH + 2.75) * 1e12;
    const double f_coll_prefactor = ReionEfficiency * (4.0 / 3.0) * M_PI / pixel_volume / M_mean * R_cubed;

    /* :451-544 per-cell criterion and state */
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < N; ix++)
      for (int iy = 0; iy < N; iy++)
        for (int iz = 0; iz < N; iz++) {
          const long ir = iz + (long)N * (iy + (long)N * ix);
          const long ip = iz + 2 * H * (iy + (long)N * ix);
          const double density_over_mean = 1.0 + (double)filtered[0][ip];
          const double f_coll_stars = (double)filtered[1][ip] / density_over_mean;
          const double f_coll_effective_bhm = bh ? (double)filtered[IBHM][ip] / density_over_mean : 0.0;
          const double rec = recomb ? (double)filtered[INREC][ip] / density_over_mean : 0.0;
          double neutral_fraction;
          if (spin)
            neutral_fraction = 1.0 - (double)filtered[IXE][ip];
          else
            neutral_fraction = 1.0;
          if (margin) {
            const double m = (f_coll_effective_bhm + f_coll_stars) * f_coll_prefactor / (neutral_fraction * (1. + rec)) - 1.0;
            if (m > margin[ir])
              margin[ir] = m;
          }
          if ((f_coll_effective_bhm + f_coll_stars) * f_coll_prefactor > neutral_fraction * (1. + rec)) {
            /* :488-506 first crossing (largest R) with recombinations */
            if ((xH[ir] > REL_TOL) && recomb) {
              Gamma12[ir] = (float)((REAL)filtered[2][ip] * (REAL)(float)Gamma_R_prefactor);
              if (bh)
                Gamma12[ir] = (float)((REAL)Gamma12[ir] + (REAL)filtered[IBHAR][ip] * (REAL)(float)Gamma_R_prefactor_BH);
              r_bubble[ir] = (float)R;
              if (flag_ReionUVBFlag == 2) {
                J_21_at_ionization[ir] = (float)((REAL)filtered[1][ip] * (REAL)J_21_aux_constant);
                if (bh)
                  J_21_at_ionization[ir] = (float)((REAL)J_21_at_ionization[ir] +
                                                   (REAL)filtered[IBHM][ip] * (REAL)J_21_aux_constant_BH);
              }
            }
            xH[ir] = 0;
          } else if (flag_last_filter_step && (xH[ir] > REL_TOL)) {
            if (spin)
              temp_kinetic_all_gas[ir] = partially_ionized_T(Tk_box[ir], xH[ir]);
            else
              temp_kinetic_all_gas[ir] = partially_ionized_T(TK * (1. + cT_ad * deltax[ip]), xH[ir]);
            xH[ir] = (float)(neutral_fraction - (f_coll_effective_bhm + f_coll_stars) * f_coll_prefactor);
            CLAMP_0_1(xH[ir]);
          }
          if ((xH[ir] < REL_TOL) && (z_in[ir] < 0)) {
            z_in[ir] = (float)redshift;
            if (flag_ReionUVBFlag == 1) {
              J_21_at_ionization[ir] = (float)((REAL)filtered[1][ip] * (REAL)J_21_aux_constant);
              if (bh)
                J_21_at_ionization[ir] = (float)((REAL)J_21_at_ionization[ir] +
                                                 (REAL)filtered[IBHM][ip] * (REAL)J_21_aux_constant_BH);
            }
          }
        }
    R /= ReionDeltaRFactor;
  }

  for (int gi = 0; gi < ngrid; gi++) {
    free(unfiltered[gi]);
    free(filtered[gi]);
  }

  if (max_radii > 0) {
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds)
      *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return n_radii;
  }

  /* :549-645 reductions and temperatures */
  double volume_weighted_global_xH = 0.0, volume_weighted_global_r_bubble = 0.0;
  double volume_weighted_global_weighted_sfr = 0.0, volume_weighted_global_temp_kinetic_all_gas = 0.0;
  double volume_weighted_global_effective_bhar = 0.0;
  double mass_weight = 0.0, mass_weighted_global_xH = 0.0, mass_weighted_global_r_bubble = 0.0;
  double mass_weighted_global_temp_kinetic_all_gas = 0.0;
  double volume_weighted_global_N_rec = 0.0, volume_weighted_global_residual_xH = 0.0;
  double volume_weighted_global_clumping_factor = 0.0, volume_weighted_global_Gamma12 = 0.0;
  double mass_weighted_global_N_rec = 0.0, mass_weighted_global_residual_xH = 0.0;
  double mass_weighted_global_clumping_factor = 0.0, mass_weighted_global_Gamma12 = 0.0;
  const double Hubble_h = p->Hubble_h;
  int rec_fail = 0;
  for (int ix = 0; ix < N; ix++)
    for (int iy = 0; iy < N; iy++)
      for (int iz = 0; iz < N; iz++) {
        const long ir = iz + (long)N * (iy + (long)N * ix);
        const long ip = iz + 2 * H * (iy + (long)N * ix);
        const double cell_xH = (double)(xH[ir]);
        volume_weighted_global_xH += cell_xH;
        volume_weighted_global_r_bubble += (double)r_bubble[ir];
        volume_weighted_global_weighted_sfr += (double)weighted_sfr[ip];
        if (bh)
          volume_weighted_global_effective_bhar += (double)g->effective_bhar[ip];
        const double density_over_mean = 1.0 + (double)deltax[ip];
        mass_weighted_global_xH += cell_xH * density_over_mean;
        mass_weighted_global_r_bubble += (double)r_bubble[ir] * density_over_mean;
        mass_weight += density_over_mean;
        if ((z_in[ir] > 0) && (xH[ir] < REL_TOL))
          temp_kinetic_all_gas[ir] = fully_ionized_T(z_in[ir], (float)redshift, deltax[ip]);
        if (spin) {
          if (temp_kinetic_all_gas[ir] < Tk_box[ir])
            temp_kinetic_all_gas[ir] = Tk_box[ir];
        } else {
          const float thistk = TK * (1. + cT_ad * deltax[ip]);
          if (temp_kinetic_all_gas[ir] < thistk)
            temp_kinetic_all_gas[ir] = thistk;
        }
        volume_weighted_global_temp_kinetic_all_gas += (double)temp_kinetic_all_gas[ir];
        mass_weighted_global_temp_kinetic_all_gas += (double)temp_kinetic_all_gas[ir] * density_over_mean;
        if (recomb) {
          /* :612-636 */
          double recombination_rate, rnh, cf;
          const double z_eff = (1. + redshift) * pow(density_over_mean, 1.0 / 3.0) - 1;
          const double temp = p->Flag_TemperatureDependentRec ? (double)temp_kinetic_all_gas[ir] : 1e4;
          if (oracle_splined_recombination(z_eff, (double)Gamma12[ir] * Hubble_h * Hubble_h, temp,
                                           &recombination_rate, &rnh, &cf) != 0) {
            rec_fail = 1;
            continue;
          }
          N_rec[ip] += (float)recombination_rate * fabs_dtdz * zstep * (1. - (float)cell_xH);
          g->residual_xH[ir] = (float)rnh;
          g->clumping_factor[ir] = (float)cf;
          volume_weighted_global_N_rec += (double)N_rec[ip];
          g->t_resp[ir] = Gamma12[ir] * Hubble_h * Hubble_h * 1e-12 + recombination_rate;
          if (g->t_resp[ir] > 0)
            g->t_resp[ir] = 1.0 / g->t_resp[ir] / SEC_PER_MEGAYEAR;
          else
            g->t_resp[ir] = 1e30;
          volume_weighted_global_residual_xH += (double)g->residual_xH[ir];
          volume_weighted_global_clumping_factor += (double)g->clumping_factor[ir];
          volume_weighted_global_Gamma12 += (double)Gamma12[ir];
          mass_weighted_global_N_rec += (double)N_rec[ip] * density_over_mean;
          mass_weighted_global_residual_xH += (double)g->residual_xH[ir] * density_over_mean;
          mass_weighted_global_clumping_factor += (double)g->clumping_factor[ir] * density_over_mean;
          mass_weighted_global_Gamma12 += (double)Gamma12[ir] * density_over_mean;
        }
      }
  if (rec_fail)
    return MXR_ERR_TABLES;

  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (seconds)
    *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

  if (out) {
    memset(out, 0, sizeof(*out));
    out->volume_weighted_global_xH = volume_weighted_global_xH / total_n_cells;
    out->volume_weighted_global_r_bubble = volume_weighted_global_r_bubble / total_n_cells;
    out->volume_weighted_global_weighted_sfr = volume_weighted_global_weighted_sfr / total_n_cells *
                                               (UnitMass_in_g / UnitTime_in_s * SEC_PER_YEAR / SOLAR_MASS);
    out->volume_weighted_global_effective_bhar = volume_weighted_global_effective_bhar / total_n_cells *
                                                 (UnitMass_in_g / UnitTime_in_s * SEC_PER_YEAR / SOLAR_MASS);
    out->volume_weighted_global_temp_kinetic_all_gas = volume_weighted_global_temp_kinetic_all_gas / total_n_cells;
    out->mass_weighted_global_xH = mass_weighted_global_xH / mass_weight;
    out->mass_weighted_global_r_bubble = mass_weighted_global_r_bubble / mass_weight;
    out->mass_weighted_global_temp_kinetic_all_gas = mass_weighted_global_temp_kinetic_all_gas / mass_weight;
    out->volume_weighted_global_N_rec = volume_weighted_global_N_rec / total_n_cells;
    out->volume_weighted_global_residual_xH = volume_weighted_global_residual_xH / total_n_cells;
    out->volume_weighted_global_clumping_factor = volume_weighted_global_clumping_factor / total_n_cells;
    out->volume_weighted_global_Gamma12 = volume_weighted_global_Gamma12 / total_n_cells;
    out->mass_weighted_global_N_rec = mass_weighted_global_N_rec / mass_weight;
    out->mass_weighted_global_residual_xH = mass_weighted_global_residual_xH / mass_weight;
    out->mass_weighted_global_clumping_factor = mass_weighted_global_clumping_factor / mass_weight;
    out->mass_weighted_global_Gamma12 = mass_weighted_global_Gamma12 / mass_weight;
    out->n_radii = n_radii;
    out->device_ms = 0.0;
  }
  return n_radii;
}

#undef CAT
#undef CAT_
