/* BuenoOrovio constant slots of the oracle's constant vector, in the order of
 * BuenoOrovio_2008Names.H (shared by models.c and the tangent-linear run). */
#ifndef ORC_BO_CONSTS_H
#define ORC_BO_CONSTS_H
enum { BO_uo, BO_uu, BO_thv, BO_thw, BO_thvm, BO_tho, BO_tv1m, BO_tv2m, BO_tvp,
       BO_tw1m, BO_tw2m, BO_kwm, BO_uwm, BO_twp, BO_tfi, BO_to1, BO_to2, BO_tso1,
       BO_tso2, BO_kso, BO_uso, BO_ts1, BO_ts2, BO_ks, BO_us, BO_tsi, BO_twinf,
       BO_winfstar, BO_NC };
#endif
