/*
 * cable_fields.h -- the per-tile fields one CABLE timestep (cbm) reads and
 * writes, as an X-macro table.
 *
 * Each entry is X(ident, "fortran name", prec, n, role):
 *   fortran name  the derived-type component in the reference
 *                 (src/offline/cable_define_types.F90:79-717);
 *   prec          F = default REAL, D = REAL(r_2) (always FP64), I = INTEGER,
 *                 B = default REAL in the reference but FP64 in every GPU mode
 *                 (the per-tile water and energy balance accumulators);
 *   n             elements per tile; the array is the Fortran (mp, ...) array,
 *                 i.e. element k of tile t sits at k*mp + t (tile fastest);
 *   role          PARAM  per-tile parameter, read only;
 *                 FORC   per-step forcing (met%...), read only;
 *                 STATE  carried from step to step (read, then written);
 *                 DIAG   written by the step (some only on a subset of tiles,
 *                        where the untouched tiles keep their previous value,
 *                        exactly as the reference's masked assignments do).
 *
 * The set was established by running the compiled reference with each field
 * perturbed before a step and recording which outputs change.
 */
#ifndef CABLE_FIELDS_H
#define CABLE_FIELDS_H

#define CABLE_FIELDS(X)                                                    \
  /* ---- vegetation parameters (veg_parameter_type) ---- */              \
  X(veg_iveg, "veg%iveg", I, 1, PARAM)                                     \
  X(veg_vlai, "veg%vlai", F, 1, PARAM)                                     \
  X(veg_canst1, "veg%canst1", F, 1, PARAM)                                 \
  X(veg_dleaf, "veg%dleaf", F, 1, PARAM)                                   \
  X(veg_ejmax, "veg%ejmax", F, 1, PARAM)                                   \
  X(veg_frac4, "veg%frac4", F, 1, PARAM)                                   \
  X(veg_hc, "veg%hc", F, 1, PARAM)                                         \
  X(veg_rp20, "veg%rp20", F, 1, PARAM)                                     \
  X(veg_shelrb, "veg%shelrb", F, 1, PARAM)                                 \
  X(veg_vegcf, "veg%vegcf", F, 1, PARAM)                                   \
  X(veg_vbeta, "veg%vbeta", F, 1, PARAM)                                   \
  X(veg_vcmax, "veg%vcmax", F, 1, PARAM)                                   \
  X(veg_xfang, "veg%xfang", F, 1, PARAM)                                   \
  X(veg_extkn, "veg%extkn", F, 1, PARAM)                                   \
  X(veg_a1gs, "veg%a1gs", F, 1, PARAM)                                     \
  X(veg_d0gs, "veg%d0gs", F, 1, PARAM)                                     \
  X(veg_alpha, "veg%alpha", F, 1, PARAM)                                   \
  X(veg_convex, "veg%convex", F, 1, PARAM)                                 \
  X(veg_cfrd, "veg%cfrd", F, 1, PARAM)                                     \
  X(veg_gswmin, "veg%gswmin", F, 1, PARAM)                                 \
  X(veg_conkc0, "veg%conkc0", F, 1, PARAM)                                 \
  X(veg_conko0, "veg%conko0", F, 1, PARAM)                                 \
  X(veg_ekc, "veg%ekc", F, 1, PARAM)                                       \
  X(veg_eko, "veg%eko", F, 1, PARAM)                                       \
  X(veg_meth, "veg%meth", F, 1, PARAM)                                     \
  X(veg_refl, "veg%refl", F, 2, PARAM)                                     \
  X(veg_taul, "veg%taul", F, 2, PARAM)                                     \
  X(veg_froot, "veg%froot", F, 6, PARAM)                                   \
  /* ---- soil parameters (soil_parameter_type) ---- */                   \
  X(soil_isoilm, "soil%isoilm", I, 1, PARAM)                               \
  X(soil_css, "soil%css", F, 1, PARAM)                                     \
  X(soil_hsbh, "soil%hsbh", F, 1, PARAM)                                   \
  X(soil_hyds, "soil%hyds", F, 1, PARAM)                                   \
  X(soil_i2bp3, "soil%i2bp3", F, 1, PARAM)                                 \
  X(soil_ibp2, "soil%ibp2", F, 1, PARAM)                                   \
  X(soil_rhosoil, "soil%rhosoil", F, 1, PARAM)                             \
  X(soil_sfc, "soil%sfc", F, 1, PARAM)                                     \
  X(soil_ssat, "soil%ssat", F, 1, PARAM)                                   \
  X(soil_swilt, "soil%swilt", F, 1, PARAM)                                 \
  X(soil_cnsd, "soil%cnsd", D, 1, PARAM)                                   \
  X(soil_pwb_min, "soil%pwb_min", D, 1, PARAM)                             \
  X(soil_albsoil, "soil%albsoil", F, 3, PARAM)                             \
  X(soil_sfc_vec, "soil%sfc_vec", D, 6, PARAM)                             \
  X(soil_swilt_vec, "soil%swilt_vec", D, 6, PARAM)                         \
  X(soil_zse_vec, "soil%zse_vec", D, 6, PARAM)                             \
  X(rough_za_uv, "rough%za_uv", F, 1, PARAM)                               \
  X(rough_za_tq, "rough%za_tq", F, 1, PARAM)                               \
  X(canopy_fes_cor, "canopy%fes_cor", D, 1, PARAM)                         \
  X(ssnow_qrecharge, "ssnow%qrecharge", D, 1, PARAM)                       \
  X(ssnow_pudsmx, "ssnow%pudsmx", F, 1, PARAM)                             \
  /* ---- forcing (met_type) ---- */                                      \
  X(met_fsd, "met%fsd", F, 2, FORC)                                        \
  X(met_fld, "met%fld", F, 1, FORC)                                        \
  X(met_tk, "met%tk", F, 1, FORC)                                          \
  X(met_tvrad, "met%tvrad", F, 1, FORC)                                    \
  X(met_pmb, "met%pmb", F, 1, FORC)                                        \
  X(met_ua, "met%ua", F, 1, FORC)                                          \
  X(met_qv, "met%qv", F, 1, FORC)                                          \
  X(met_precip, "met%precip", F, 1, FORC)                                  \
  X(met_precip_sn, "met%precip_sn", F, 1, FORC)                            \
  X(met_ca, "met%ca", F, 1, FORC)                                          \
  X(met_coszen, "met%coszen", F, 1, FORC)                                  \
  X(met_doy, "met%doy", F, 1, FORC)                                        \
  /* met%tvair is set to met%tk by the forcing reader and rewritten by    \
     the canopy scheme: forcing in, diagnostic out */                     \
  X(met_tvair, "met%tvair", F, 1, FORC)                                    \
  /* ---- prognostic state ---- */                                        \
  X(ssnow_isflag, "ssnow%isflag", I, 1, STATE)                             \
  X(ssnow_pudsto, "ssnow%pudsto", F, 1, STATE)                             \
  X(ssnow_rtsoil, "ssnow%rtsoil", F, 1, STATE)                             \
  X(ssnow_wbtot1, "ssnow%wbtot1", F, 1, STATE)                             \
  X(ssnow_wbtot2, "ssnow%wbtot2", F, 1, STATE)                             \
  X(ssnow_wb_lake, "ssnow%wb_lake", F, 1, STATE)                           \
  X(ssnow_owetfac, "ssnow%owetfac", F, 1, STATE)                           \
  X(ssnow_snage, "ssnow%snage", F, 1, STATE)                               \
  X(ssnow_snowd, "ssnow%snowd", F, 1, STATE)                               \
  X(ssnow_ssdnn, "ssnow%ssdnn", F, 1, STATE)                               \
  X(ssnow_tss, "ssnow%tss", F, 1, STATE)                                   \
  X(ssnow_smass, "ssnow%smass", F, 3, STATE)                               \
  X(ssnow_ssdn, "ssnow%ssdn", F, 3, STATE)                                 \
  X(ssnow_tgg, "ssnow%tgg", F, 6, STATE)                                   \
  X(ssnow_tggsn, "ssnow%tggsn", F, 3, STATE)                               \
  X(ssnow_gammzz, "ssnow%gammzz", D, 6, STATE)                             \
  X(ssnow_wb, "ssnow%wb", D, 6, STATE)                                     \
  X(ssnow_wbice, "ssnow%wbice", D, 6, STATE)                               \
  X(ssnow_wbliq, "ssnow%wbliq", D, 6, STATE)                               \
  X(canopy_cansto, "canopy%cansto", F, 1, STATE)                           \
  X(bgc_cplant, "bgc%cplant", F, 3, STATE)                                 \
  X(bgc_csoil, "bgc%csoil", F, 2, STATE)                                   \
  X(bal_ebal_tot, "bal%ebal_tot", B, 1, STATE)                             \
  X(bal_wbal_tot, "bal%wbal_tot", B, 1, STATE)                             \
  X(bal_precip_tot, "bal%precip_tot", B, 1, STATE)                         \
  X(bal_rnoff_tot, "bal%rnoff_tot", B, 1, STATE)                           \
  X(bal_evap_tot, "bal%evap_tot", B, 1, STATE)                             \
  X(bal_radbalsum, "bal%radbalsum", B, 1, STATE)                           \
  /* total soil water at the previous balance check (the SAVEd owb of     \
     src/offline/cable_checks.F90:496-513), one value per tile */          \
  X(drv_owb, "drv%owb", D, 1, STATE)                                       \
  /* the offline driver's running carbon sums, each flux times the step   \
     length (sumcflux, src/offline/casa_cable.F90:482-560, called at      \
     src/offline/cable_serial.F90:713-715); dsumrs, sumxrp and sumxrs are \
     not updated there */                                                  \
  X(sum_flux_sumpn, "sum_flux%sumpn", F, 1, STATE)                         \
  X(sum_flux_sumrp, "sum_flux%sumrp", F, 1, STATE)                         \
  X(sum_flux_sumrpw, "sum_flux%sumrpw", F, 1, STATE)                       \
  X(sum_flux_sumrpr, "sum_flux%sumrpr", F, 1, STATE)                       \
  X(sum_flux_sumrs, "sum_flux%sumrs", F, 1, STATE)                         \
  X(sum_flux_sumrd, "sum_flux%sumrd", F, 1, STATE)                         \
  X(sum_flux_dsumpn, "sum_flux%dsumpn", F, 1, STATE)                       \
  X(sum_flux_dsumrp, "sum_flux%dsumrp", F, 1, STATE)                       \
  X(sum_flux_dsumrd, "sum_flux%dsumrd", F, 1, STATE)                       \
  /* ---- diagnostics ---- */                                             \
  X(met_qvair, "met%qvair", F, 1, DIAG)                                    \
  X(met_dva, "met%dva", F, 1, DIAG)                                        \
  X(met_ofsd, "met%ofsd", F, 1, DIAG)                                      \
  X(air_rho, "air%rho", F, 1, DIAG)                                        \
  X(air_volm, "air%volm", F, 1, DIAG)                                      \
  X(air_rlam, "air%rlam", F, 1, DIAG)                                      \
  X(air_qsat, "air%qsat", F, 1, DIAG)                                      \
  X(air_epsi, "air%epsi", F, 1, DIAG)                                      \
  X(air_visc, "air%visc", F, 1, DIAG)                                      \
  X(air_psyc, "air%psyc", F, 1, DIAG)                                      \
  X(air_dsatdk, "air%dsatdk", F, 1, DIAG)                                  \
  X(air_cmolar, "air%cmolar", F, 1, DIAG)                                  \
  X(rough_hruff, "rough%hruff", F, 1, DIAG)                                \
  X(rough_z0soil, "rough%z0soil", F, 1, DIAG)                              \
  X(rough_z0soilsn, "rough%z0soilsn", F, 1, DIAG)                          \
  X(rough_z0m, "rough%z0m", F, 1, DIAG)                                    \
  X(rough_rt0us, "rough%rt0us", F, 1, DIAG)                                \
  X(rough_disp, "rough%disp", F, 1, DIAG)                                  \
  X(rough_zref_uv, "rough%zref_uv", F, 1, DIAG)                            \
  X(rough_zref_tq, "rough%zref_tq", F, 1, DIAG)                            \
  X(rough_zruffs, "rough%zruffs", F, 1, DIAG)                              \
  X(rough_rt1usa, "rough%rt1usa", F, 1, DIAG)                              \
  X(rough_rt1usb, "rough%rt1usb", F, 1, DIAG)                              \
  X(rough_usuh, "rough%usuh", F, 1, DIAG)                                  \
  X(rough_coexp, "rough%coexp", F, 1, DIAG)                                \
  X(rough_term2, "rough%term2", F, 1, DIAG)                                \
  X(rough_term3, "rough%term3", F, 1, DIAG)                                \
  X(rough_term5, "rough%term5", F, 1, DIAG)                                \
  X(rough_term6, "rough%term6", F, 1, DIAG)                                \
  X(rough_term6a, "rough%term6a", F, 1, DIAG)                              \
  X(rough_rt1, "rough%rt1", F, 1, DIAG)                                    \
  X(rad_extkb, "rad%extkb", F, 1, DIAG)                                    \
  X(rad_extkd, "rad%extkd", F, 1, DIAG)                                    \
  X(rad_extkbm, "rad%extkbm", F, 3, DIAG)                                  \
  X(rad_extkdm, "rad%extkdm", F, 3, DIAG)                                  \
  X(rad_fbeam, "rad%fbeam", F, 3, DIAG)                                    \
  X(rad_albedo, "rad%albedo", F, 3, DIAG)                                  \
  X(rad_albedo_t, "rad%albedo_t", F, 1, DIAG)                              \
  X(rad_rhocbm, "rad%rhocbm", F, 3, DIAG)                                  \
  X(rad_rhocdf, "rad%rhocdf", F, 3, DIAG)                                  \
  X(rad_cexpkbm, "rad%cexpkbm", F, 2, DIAG)                                \
  X(rad_cexpkdm, "rad%cexpkdm", F, 2, DIAG)                                \
  X(rad_reffbm, "rad%reffbm", F, 3, DIAG)                                  \
  X(rad_reffdf, "rad%reffdf", F, 3, DIAG)                                  \
  X(rad_transd, "rad%transd", F, 1, DIAG)                                  \
  X(rad_transb, "rad%transb", F, 1, DIAG)                                  \
  X(rad_flws, "rad%flws", F, 1, DIAG)                                      \
  X(rad_gradis, "rad%gradis", F, 2, DIAG)                                  \
  X(rad_qcan, "rad%qcan", F, 6, DIAG)                                      \
  X(rad_qssabs, "rad%qssabs", F, 1, DIAG)                                  \
  X(rad_scalex, "rad%scalex", F, 2, DIAG)                                  \
  X(rad_fvlai, "rad%fvlai", F, 2, DIAG)                                    \
  X(rad_rniso, "rad%rniso", F, 2, DIAG)                                    \
  X(rad_lwabv, "rad%lwabv", F, 1, DIAG)                                    \
  X(rad_swnet, "rad%swnet", F, 1, DIAG)                                    \
  X(rad_lwnet, "rad%lwnet", F, 1, DIAG)                                    \
  X(rad_rnet, "rad%rnet", F, 1, DIAG)                                      \
  X(rad_trad, "rad%trad", F, 1, DIAG)                                      \
  X(canopy_vlaiw, "canopy%vlaiw", F, 1, DIAG)                              \
  X(canopy_rghlai, "canopy%rghlai", F, 1, DIAG)                            \
  X(canopy_oldcansto, "canopy%oldcansto", F, 1, DIAG)                      \
  X(canopy_wcint, "canopy%wcint", F, 1, DIAG)                              \
  X(canopy_through, "canopy%through", F, 1, DIAG)                          \
  X(canopy_fwet, "canopy%fwet", F, 1, DIAG)                                \
  X(canopy_fevw_pot, "canopy%fevw_pot", F, 1, DIAG)                        \
  X(canopy_gswx, "canopy%gswx", F, 2, DIAG)                                \
  X(canopy_tv, "canopy%tv", F, 1, DIAG)                                    \
  X(canopy_fwsoil, "canopy%fwsoil", D, 1, DIAG)                            \
  X(canopy_kthlitt, "canopy%kthlitt", D, 1, DIAG)                          \
  X(canopy_dvlitt, "canopy%dvlitt", D, 1, DIAG)                            \
  X(canopy_fes, "canopy%fes", D, 1, DIAG)                                  \
  X(canopy_fess, "canopy%fess", D, 1, DIAG)                                \
  X(canopy_fesp, "canopy%fesp", D, 1, DIAG)                                \
  X(canopy_fhs_cor, "canopy%fhs_cor", F, 1, DIAG)                          \
  X(canopy_fns_cor, "canopy%fns_cor", F, 1, DIAG)                          \
  X(canopy_ga_cor, "canopy%ga_cor", F, 1, DIAG)                            \
  X(canopy_zetar, "canopy%zetar", F, 4, DIAG)                              \
  X(canopy_zetash, "canopy%zetash", F, 4, DIAG)                            \
  X(canopy_us, "canopy%us", F, 1, DIAG)                                    \
  X(canopy_fevc, "canopy%fevc", D, 1, DIAG)                                \
  X(canopy_fevw, "canopy%fevw", F, 1, DIAG)                                \
  X(canopy_fhvw, "canopy%fhvw", F, 1, DIAG)                                \
  X(canopy_fev, "canopy%fev", F, 1, DIAG)                                  \
  X(canopy_fhv, "canopy%fhv", F, 1, DIAG)                                  \
  X(canopy_fnv, "canopy%fnv", F, 1, DIAG)                                  \
  X(canopy_fns, "canopy%fns", F, 1, DIAG)                                  \
  X(canopy_fhs, "canopy%fhs", F, 1, DIAG)                                  \
  X(canopy_ga, "canopy%ga", F, 1, DIAG)                                    \
  X(canopy_fe, "canopy%fe", F, 1, DIAG)                                    \
  X(canopy_fh, "canopy%fh", F, 1, DIAG)                                    \
  X(canopy_rnet, "canopy%rnet", F, 1, DIAG)                                \
  X(canopy_rniso, "canopy%rniso", F, 1, DIAG)                              \
  X(canopy_epot, "canopy%epot", F, 1, DIAG)                                \
  X(canopy_wetfac_cs, "canopy%wetfac_cs", F, 1, DIAG)                      \
  X(canopy_cduv, "canopy%cduv", F, 1, DIAG)                                \
  X(canopy_gswx_t, "canopy%gswx_t", F, 1, DIAG)                            \
  X(canopy_cdtq, "canopy%cdtq", F, 1, DIAG)                                \
  X(canopy_tscrn, "canopy%tscrn", F, 1, DIAG)                              \
  X(canopy_qscrn, "canopy%qscrn", F, 1, DIAG)                              \
  X(canopy_qmom, "canopy%qmom", F, 1, DIAG)                                \
  X(canopy_dewmm, "canopy%dewmm", F, 1, DIAG)                              \
  X(canopy_spill, "canopy%spill", F, 1, DIAG)                              \
  X(canopy_precis, "canopy%precis", F, 1, DIAG)                            \
  X(canopy_delwc, "canopy%delwc", F, 1, DIAG)                              \
  X(canopy_dgdtg, "canopy%dgdtg", D, 1, DIAG)                              \
  X(canopy_frday, "canopy%frday", F, 1, DIAG)                              \
  X(canopy_fpn, "canopy%fpn", F, 1, DIAG)                                  \
  X(canopy_sghflux, "canopy%sghflux", F, 1, DIAG)                          \
  X(canopy_ghflux, "canopy%ghflux", F, 1, DIAG)                            \
  X(canopy_segg, "canopy%segg", F, 1, DIAG)                                \
  X(canopy_frp, "canopy%frp", F, 1, DIAG)                                  \
  X(canopy_frpw, "canopy%frpw", F, 1, DIAG)                                \
  X(canopy_frpr, "canopy%frpr", F, 1, DIAG)                                \
  X(canopy_frs, "canopy%frs", F, 1, DIAG)                                  \
  X(canopy_fnpp, "canopy%fnpp", F, 1, DIAG)                                \
  X(canopy_fgpp, "canopy%fgpp", F, 1, DIAG)                                \
  X(canopy_fnee, "canopy%fnee", F, 1, DIAG)                                \
  X(canopy_fra, "canopy%fra", F, 1, DIAG)                                  \
  X(ssnow_otss_0, "ssnow%otss_0", F, 1, DIAG)                              \
  X(ssnow_otss, "ssnow%otss", F, 1, DIAG)                                  \
  X(ssnow_wetfac, "ssnow%wetfac", F, 1, DIAG)                              \
  X(ssnow_satfrac, "ssnow%satfrac", D, 1, DIAG)                            \
  X(ssnow_rh_srf, "ssnow%rh_srf", D, 1, DIAG)                              \
  X(ssnow_evapfbl, "ssnow%evapfbl", D, 6, DIAG)                            \
  X(ssnow_rex, "ssnow%rex", D, 6, DIAG)                                    \
  X(ssnow_potev, "ssnow%potev", F, 1, DIAG)                                \
  X(ssnow_qstss, "ssnow%qstss", F, 1, DIAG)                                \
  X(ssnow_cls, "ssnow%cls", F, 1, DIAG)                                    \
  X(ssnow_dfn_dtg, "ssnow%dfn_dtg", F, 1, DIAG)                            \
  X(ssnow_dfh_dtg, "ssnow%dfh_dtg", F, 1, DIAG)                            \
  X(ssnow_dfe_ddq, "ssnow%dfe_ddq", F, 1, DIAG)                            \
  X(ssnow_ddq_dtg, "ssnow%ddq_dtg", F, 1, DIAG)                            \
  X(ssnow_dfe_dtg, "ssnow%dfe_dtg", F, 1, DIAG)                            \
  X(ssnow_tggav, "ssnow%tggav", F, 1, DIAG)                                \
  X(ssnow_t_snwlr, "ssnow%t_snwlr", F, 1, DIAG)                            \
  X(ssnow_fwtop1, "ssnow%fwtop1", F, 1, DIAG)                              \
  X(ssnow_fwtop2, "ssnow%fwtop2", F, 1, DIAG)                              \
  X(ssnow_fwtop3, "ssnow%fwtop3", F, 1, DIAG)                              \
  X(ssnow_runoff, "ssnow%runoff", F, 1, DIAG)                              \
  X(ssnow_rnof1, "ssnow%rnof1", F, 1, DIAG)                                \
  X(ssnow_rnof2, "ssnow%rnof2", F, 1, DIAG)                                \
  X(ssnow_smelt, "ssnow%smelt", F, 1, DIAG)                                \
  X(ssnow_dtmlt, "ssnow%dtmlt", F, 3, DIAG)                                \
  X(ssnow_osnowd, "ssnow%osnowd", F, 1, DIAG)                              \
  X(ssnow_wblf, "ssnow%wblf", D, 6, DIAG)                                  \
  X(ssnow_wbfice, "ssnow%wbfice", D, 6, DIAG)                              \
  X(ssnow_sdepth, "ssnow%sdepth", F, 3, DIAG)                              \
  X(ssnow_sconds, "ssnow%sconds", F, 3, DIAG)                              \
  X(ssnow_evapsn, "ssnow%evapsn", F, 1, DIAG)                              \
  X(ssnow_sinfil, "ssnow%sinfil", F, 1, DIAG)                              \
  X(ssnow_totsdepth, "ssnow%totsdepth", F, 1, DIAG)                        \
  X(ssnow_wbtot, "ssnow%wbtot", D, 1, DIAG)                                \
  X(ssnow_deltss, "ssnow%deltss", F, 1, DIAG)                              \
  X(ssnow_albsoilsn, "ssnow%albsoilsn", F, 3, DIAG)                        \
  X(soil_heat_cap_lower_limit, "soil%heat_cap_lower_limit", F, 6, DIAG)    \
  X(bal_drybal, "bal%drybal", F, 1, DIAG)                                  \
  X(bal_wetbal, "bal%wetbal", F, 1, DIAG)                                  \
  X(bal_wbal, "bal%wbal", F, 1, DIAG)                                      \
  X(bal_radbal, "bal%radbal", F, 1, DIAG)                                  \
  X(bal_ebalsoil, "bal%ebalsoil", F, 1, DIAG)                              \
  X(bal_ebalveg, "bal%ebalveg", F, 1, DIAG)                                \
  X(bal_ebal, "bal%ebal", F, 1, DIAG)

/* Field identifiers: CABLE_F_<ident>, in table order. */
enum cable_field_id {
#define CABLE_FIELD_ENUM_(id, name, prec, n, role) CABLE_F_##id,
  CABLE_FIELDS(CABLE_FIELD_ENUM_)
#undef CABLE_FIELD_ENUM_
  CABLE_NFIELDS
};

#endif /* CABLE_FIELDS_H */
