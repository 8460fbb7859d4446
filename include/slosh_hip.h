/* slosh_hip.h -- C ABI of the MI355X (gfx950) SLOSH time-stepping core.
 *
 * The reference (NOAA-MDL/slosh-test v4.23) runs one storm per process: its C
 * driver calls the Fortran step routines TMSTPC1 / TMSTPC2 / TMSTPM once per
 * time step (sorc/slosh.fd/slosh2.c:1400, 1540, 1542) on state held in Fortran
 * COMMON blocks.  This library replaces that path.  Its handle API owns the
 * basin statics and any number of storm runs on one basin, resident in HBM and
 * stepped together by hand-written HIP kernels, in FP64 (the reference's
 * default build) or FP32.
 *
 * Plain pointers and sizes only; no torch or HIP types cross this boundary.
 * Functions return 0 on success or a negative SLG_E* code, with a message in
 * err[errlen] where an err argument is present.
 */
#ifndef SLOSH_HIP_H
#define SLOSH_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SLG_EINPUT (-1)  /* bad or unreadable input file / argument */
#define SLG_EUNSUP (-2)  /* basin or track feature this build does not implement */
#define SLG_EHIP (-3)    /* HIP runtime error (device missing, out of memory, ...) */
#define SLG_ELIMIT (-4)  /* a reference hard limit was reached (the reference STOPs) */
#define SLG_MAX_POINTS 1048576 /* squares of one point series (slg_batch_rex_points) */
#define SLG_MAX_STORMS 65535   /* runs of one batch (a launch grid dimension holds them) */

typedef struct slg_basin slg_basin;
typedef struct slg_batch slg_batch;

/* Load a basin: the dta file (read as CRDRD1/CRDRD2 do, crdrd1.f:102-213,
 * crdrd2.f:65-489) and its ?basins.dta grid line (memSetLLx, savellx.c:704-765),
 * then derive masks, loop bounds, metrics, bottom-stress tables and constants
 * (INITLZ, initlz.f:61-84).  Replaces the basin half of INITAL (inital.f:1-277).
 * bnt_line may be NULL if slg_basin_set_ylt supplies the cell latitudes. */
int slg_basin_open(const char *dta_path, const char *bnt_line, slg_basin **out, char *err,
                   int errlen);
/* As slg_basin_open, with options: SLG_OPEN_FULL_DEPTH takes the depths as read
 * (the -fullDepth option, INITAL's MFULLDEP: crdrd2.f:441-444) instead of
 * capping them at 600 ft. */
#define SLG_OPEN_FULL_DEPTH 1
int slg_basin_open_opts(const char *dta_path, const char *bnt_line, int flags, slg_basin **out, char *err,
                        int errlen);
/* Cell latitudes YLT (double, [jmxb][imxb], I fastest) from the caller, as the
 * reference's C side fills COMMON /LLXFLE/ before INITAL (savellx.c:757-762). */
int slg_basin_set_ylt(slg_basin *b, const double *ylt);
int slg_basin_dims(const slg_basin *b, int *imxb, int *jmxb, int *nbcpt);
/* Counts of the basin's sub-grid 1-D flow sides as CRDRD2 reads them (COMMON
 * /FLWCPT/ NSQRS, NSQRW, NSQRWC; crdrd2.f:104-357): 1-D flow sides, plus
 * raised weirs, plus cuts.  Any pointer may be NULL. */
int slg_basin_sides(const slg_basin *b, int *nsqrs, int *nsqrw, int *nsqrwc);
void slg_basin_close(slg_basin *b);
/* Static field of the basin as double [jmxb][imxb]: "ZB", "ZBM" (COMMON /DUMB5/,
 * depsfc.f:83-138) or "YLT" (/LLXFLE/). */
int slg_basin_field(const slg_basin *b, const char *field, double *dst);

/* Open nstorm runs (1 .. SLG_MAX_STORMS) on one basin, one trk file each
 * (INPUTN2, STERON, INTVAL, SETCMP, INTLHT: inputn2.f, steron.f, intval.f,
 * setcmp.f, intlht.f).  Each run's storm is walked once on the host here
 * (track, STERON, SETCMP, STMVAL with its FORCE1 table refreshes) into a
 * compact storm program; slg_batch_upload expands the programs into the
 * per-step storm values on the device (see SLG_SCHED_HOST). */
int slg_batch_open(slg_basin *b, int nstorm, const char *const *trk_paths, slg_batch **out,
                   char *err, int errlen);
/* As slg_batch_open, choosing the arithmetic of the time step: 32 (FP32) or 64
 * (FP64, the precision of the reference's default build, -fdefault-real-8:
 * Makefile:15-21; its results reproduce that build bit for bit). */
int slg_batch_open_prec(slg_basin *b, int nstorm, const char *const *trk_paths, int precision,
                        slg_batch **out, char *err, int errlen);
/* As slg_batch_open_prec, with flags.  SLG_SCHED_HOST: compute and keep the
 * per-step storm values (TTIMER's BT, the storm position and the forcing
 * constants each STMVAL call leaves for the next step: ttimer.f:33-42,
 * stmval.f:95-236) on the host, as the reference's driver loop produces them,
 * and upload them; slg_batch_static_height needs this.  Without it they are
 * generated on the device at upload.  The values are identical either way. */
#define SLG_SCHED_HOST 1
int slg_batch_open_opts(slg_basin *b, int nstorm, const char *const *trk_paths, int precision, int flags,
                        slg_batch **out, char *err, int errlen);
/* A track in memory: what a FileVersion=2013 trk file holds (inputn2.f:110-605)
 * without the file.  Rows as the file's rows: hour, latitude (deg N), longitude
 * (deg E, negative west), pressure deficit (mb; PressureType 0) and RMW (mi);
 * a row whose hour is 999 ends the track.  wind_factor and rmw_factor: 1 when
 * a file gives none.  spline: InterpolationType Spline (1) or Linear (0). */
typedef struct slg_track {
   int year, month, day, hour_of_day, minute;   /* DateTime */
   int date_hour, landfall_hour, begin_hour, end_hour;
   double ocean_datum, lake_datum;              /* ft */
   double wind_factor, rmw_factor;
   int spline;
   int npts;
   const double *hour, *lat, *lon, *delp, *rmw;  /* [npts] */
} slg_track;
/* As slg_batch_open_opts with the tracks in memory (no per-storm files). */
int slg_batch_open_tracks(slg_basin *b, int nstorm, const slg_track *tracks, int precision, int flags,
                          slg_batch **out, char *err, int errlen);
/* MEOW track sets in memory.  The seed points are the landfall points of
 * parallel tracks heading dir across the basin, apart miles apart, covering
 * the basin's extent across the heading widened by r34 on both sides, as the
 * reference's MEOW generator places them (conMeowGenCmd,
 * sorc/rexout.cd/convert.c:1181-1288; xyflag 0 steps along the basin's
 * tangent plane, 1 along great circles).  Each track runs straight through
 * its seed at speed mi/h on the heading, hours 1..100 as the GUI's MEOW tracks
 * (CreateMEOWTrk, gui/tclsrc/meow.tcl:290-344), passing the seed at
 * landfall_hour, its points stepped along great circles (Earth_Jump,
 * convert.c:1127-1151), with constant delp and rmw.  lat/lon of the central
 * point and of the seeds: degrees, longitude positive west as the basin's
 * ?basins.dta line holds it.  Needs a basin opened with its bnt line. */
typedef struct slg_meow {
   int xyflag;
   double lat, lon;           /* central seed point */
   double dir;                /* heading of motion, deg from north */
   double apart, r34;         /* mi */
   int num;                   /* at most num tracks (-1: as many as the extent holds) */
   double speed;              /* mi/h */
   double delp, rmw;          /* mb, mi */
   int landfall_hour, begin_hour, end_hour;   /* within 1..100 */
   double ocean_datum, lake_datum;            /* ft */
} slg_meow;
#define SLG_MEOW_HOURS 100
/* Seed points into idx (the reference's numbering: 0, 1, 2, ..., then -1,
 * -2, ...), lat, lon [cap]; returns their count or a negative SLG_E*. */
int slg_meow_seeds(const slg_basin *b, const slg_meow *m, int cap, int *idx, double *lat, double *lon,
                   char *err, int errlen);
/* One track per seed, in seed order, into tracks[cap]; rows must hold
 * cap * 5 * SLG_MEOW_HOURS doubles, which the tracks point into (hour, lat,
 * lon east positive, delp, rmw).  Returns the count or a negative SLG_E*. */
int slg_meow_tracks(const slg_basin *b, const slg_meow *m, int cap, slg_track *tracks, double *rows, char *err,
                    int errlen);
/* The per-step storm values of one run as its steps read them,
 * [nsteps][SLG_NPAR] widened to double: from the device after
 * slg_batch_upload (from_device != 0), or from the host schedule of a batch
 * opened with SLG_SCHED_HOST. */
#define SLG_NPAR 26
int slg_batch_schedule(const slg_batch *x, int storm, int from_device, double *dst, char *err, int errlen);
int slg_batch_precision(const slg_batch *x);
/* Steps of one run (MHALT after any time-step change, intval.f:68). */
int slg_batch_nsteps(const slg_batch *x, int storm);
/* Allocate device state, upload statics, schedules and initial water. */
int slg_batch_upload(slg_batch *x, char *err, int errlen);
/* Advance every run by n time steps (asynchronous; one step = the reference's
 * TMSTPC1 + TMSTPC2 + TMSTPM, slosh2.c:1400-1545). */
int slg_batch_step(slg_batch *x, int n, char *err, int errlen);
int slg_batch_sync(slg_batch *x, char *err, int errlen);
/* Kernel slots of one step, for slg_batch_step_timed. */
#define SLG_K_CONTTY 0   /* CONTTY Jacobi pass */
#define SLG_K_REPLAY 1   /* CONTTY drying-cascade replay */
#define SLG_K_PFLAG 2    /* SMPT2G momentum points */
#define SLG_K_HCLASS 3   /* SMPT2G height-point classes */
#define SLG_K_SMOOTH 4   /* FLT2G + BDRYHT */
#define SLG_K_HMXSV 5    /* HMXSV envelope */
#define SLG_K_FILTER 6   /* FILTER / XFILTER */
#define SLG_K_MOMNTM 7   /* MOMNTM + FRCPNT + MNTMBD */
#define SLG_K_STEPB 8    /* fused path: SMPT2G..MOMNTM in one tiled kernel */
#define SLG_NKERN 9
/* Launch paths.  FUSED (default): CONTTY, the cascade replay, then one kernel
 * for SMPT2G/FLT2G/BDRYHT/HMXSV/FILTER/MOMNTM on LDS tiles of 512 points,
 * (512 / tile_rows) x tile_rows.  STAGED: one kernel per routine (slots 0-7).  Results are identical.
 * Basins with sub-grid 1-D flow sides (1-D points, weirs, cuts: FLW1DM_C,
 * flw1dm_block_c.f:83-435, and FLW1DM_M, flw1dm_block_m.f:1-403, plus
 * SMPT2G's side pass, smpt2g.f:101-123) or channels run those passes in
 * kernels of their own between SMPT2G/FLT2G/BDRYHT and HMXSV on either path. */
#define SLG_PATH_FUSED 0
#define SLG_PATH_STAGED 1
/* Choose the launch path and the fused tile height: 8, 16 or 32 rows, for
 * tiles of 64 x 8, 32 x 16 or 16 x 32 points. */
int slg_batch_set_path(slg_batch *x, int path, int tile_rows);
/* Replay the fused step loop as hipGraphs of steps_per_graph steps (0: off).
 * The captured steps let every storm's own step flags decide, so one graph
 * serves the whole run. */
int slg_batch_set_graph(slg_batch *x, int steps_per_graph);
/* Step the fused path's batch in nstreams parts of storms (1-4, at most one
 * per storm; 1 is the default), each on its own HIP stream; call after
 * slg_batch_upload.  The parts' kernels overlap on the device: one part's
 * CONTTY cascade replay (a serial walk per storm, contty.f:146-213) runs
 * beside the other parts' stencil kernels.  Fields are identical either way.
 * hipGraph replay uses one stream. */
int slg_batch_set_streams(slg_batch *x, int nstreams);
int slg_batch_streams(const slg_batch *x);
/* Online autotuner: times fused tiles of 8/16/32 rows and the staged path on
 * consecutive chunks of steps_per_trial real steps (all give identical fields)
 * and keeps the fastest; ms_out[4] (may be NULL) gets each candidate's time per
 * step.  With the fused path it then times storm groups of 1, 8 and 16 against
 * the default 4 (slg_batch_storm_group gives the choice). */
int slg_batch_autotune(slg_batch *x, int steps_per_trial, int *rows_out, int *path_out, double *ms_out,
                       char *err, int errlen);
/* slg_batch_step with HIP events around every kernel on the batch's stream
 * (the host waits once every 64 steps): ms[k] += kernel k time, launches[k] += 1. */
int slg_batch_step_timed(slg_batch *x, int n, double *ms, int *launches, char *err, int errlen);
int slg_batch_current_step(const slg_batch *x);
/* Storms whose tiles follow each other in the fused kernels' workgroup order
 * (1 .. 64; default 4; call after slg_batch_upload): the tiles of a group of
 * storms go out tile by tile, so a tile's static basin data (ZB, ZBM, flags,
 * Coriolis factor), shared by every storm, is read from HBM once per group and
 * from L2 by the others (the CONTTY pass keeps storm order).  Fields are
 * identical for any value; the autotuner picks one of 1, 4, 8 and 16. */
int slg_batch_set_storm_group(slg_batch *x, int sg);
int slg_batch_storm_group(const slg_batch *x);
/* For checking the CONTTY drying-cascade replay (contty.f:146-213, the serial
 * back-step sweep) on ordinary runs: evcap (1 .. 4096) drying squares are
 * listed per storm before the replay falls back to visiting every square (the
 * same results), and budget (> 0; 0 = the default bound of 64 visits per
 * square) visits before a run is stopped and flagged.  A flagged run's fields
 * are refused with SLG_ELIMIT by slg_batch_get, slg_batch_envelopes,
 * slg_batch_rex_frames, slg_batch_rex_points and slg_batch_maxstat_env (the
 * reference STOPs on its hard limits, e.g. smpt2g.f:34-38).  Call after
 * slg_batch_upload. */
int slg_batch_set_replay_limits(slg_batch *x, int evcap, long budget);
/* Copy one run's field to the host as double [jmxb][imxb]: "HB", "UB", "VB" or
 * "HMX" (COMMON /DUMB7/ and /DUMB10/, parm.for:39-45); "HB0" the initial water
 * (INTLHT, intlht.f:53-185); "HBS" HB of the last step after HMXSV and before
 * FILTER, the state TMSTPC2 copies to the C driver's HHB (tmstpc2.f:70-85;
 * kept on the staged path, and on either path when the step did not filter);
 * "FLWSQR" the transport across each sub-grid 1-D flow side (COMMON /ACT1D/,
 * parm.for:58; nsqrwc values into dst). */
int slg_batch_get(slg_batch *x, int storm, const char *field, double *dst, char *err, int errlen);
/* Values TMSTPM returns to the C driver after step n (tmstpm.f:59-70; 0 =
 * after INITAL): ITIME, MHALT, DELT, storm lat, storm lon (XY2LLX), DELP,
 * size, speed, direction.  Available for batches of up to 8 runs. */
int slg_batch_track(const slg_batch *x, int storm, int n, double *out9);
/* Run start date INITAL returns (inital.f:204-249): year, month, day, hour, min. */
int slg_batch_start_date(const slg_batch *x, int storm, int *out5);
/* Per-run counters since upload: [0] corner-smoothing momentum points, [1] dry
 * corners, [2] smoothing list 1, [3] smoothing list 2, [4] steps that replayed
 * a drying cascade. */
int slg_batch_stats(slg_batch *x, int storm, int *out8, char *err, int errlen);
void slg_batch_close(slg_batch *x);

/* Envelopes of every run as EnvSave quantises them (slosh2.c:1611-1671: int16
 * tenths of a foot, 999 dry, 0 on the last row and column), env[nstorm][jmxb]
 * [imxb], and the MOM of the batch, the maximum of MEOWs (gui/tclsrc/meowgen.tcl:
 * 34-110: per cell the largest envelope below 50 ft, 999 where none),
 * mom[jmxb][imxb]; computed on the device, either output may be NULL. */
int slg_batch_envelopes(slg_batch *x, short *env, short *mom, char *err, int errlen);
/* As slg_batch_envelopes, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops that forbid hipMalloc).  The device staging is
 * a fixed 32 MiB scratch array of the library, one per GPU, held until the
 * process ends; storms go through it in chunks.  SLG_ELIMIT if one run's
 * envelope and the MOM do not fit it (basins above 8 million points).  Calls
 * that overlap on one GPU share the scratch and race. */
int slg_batch_envelopes_noalloc(slg_batch *x, short *env, short *mom, char *err, int errlen);
/* Datums a track carries for the env trailer (ReadTrkFile, slosh2.c:1885-1975;
 * both track formats),
 * after the surge-only InitWater_TideModeOverride (slosh2.c:408-482, 2238). */
int slg_track_datums(const char *trk_path, float *ht1, float *ht2);
/* The env comment of a run (ReadTrkFile, slosh2.c:1884-1975): a FileVersion=2013
 * file's ExtraInfo= value or "FileVersion=2013"; a 1992-format file's two title
 * lines, each cut or blank-padded to 80 columns.  out: at least 161 bytes. */
int slg_track_env_comment(const char *trk_path, char *out, int cap);

/* ---- rex v2 files of batched runs ---------------------------------------
 * The frames' bit streams are encoded on the device: the basin loop of
 * RexSaveStep / RexSaveEnv (rex.c:453-499, 554-620) with the run-length
 * records of Stuff2_xxx (pack.c:166-212), for storms s0 .. s0+ns-1 at once.
 * field "HHB": HHB = min(HB, HMX) of the last step, the array TMSTPC2 hands
 * the C driver for a rex frame (tmstpc2.f:74-85; needs HB before FILTER, so
 * the last step ran on the staged path or did not filter); "HMX": the
 * envelope, the last frame RexSaveEnv writes; "HBMAX": the windowed maxima
 * (slg_batch_set_maxstat), the frames of the -f_stat mode (slosh2.c:2460).  Frame k goes to buf + k*cap
 * and its length to nbytes[k]; cap = (18*(imxb-1)*(jmxb-1) + 7) / 8 always
 * suffices (SLG_ELIMIT if a frame does not fit). */
int slg_batch_rex_frames(slg_batch *x, const char *field, int s0, int ns, unsigned char *buf, long cap,
                         int *nbytes, char *err, int errlen);
/* P-Surge windowed maxima of every run (slosh2.c:1227-1316; the reference's
 * -f_stat mode): on = 1 keeps, in double as the C driver does, the maximum of
 * HB after each step on every square, starting from -ZB (MaxStatInitGrid);
 * on = 0 drops them.  Call after slg_batch_upload.  reset starts a new window
 * (MaxStatInitGrid flag 1); env gives the EnvSave values of the window's
 * maxima (slosh2.c:1611-1671, as PerformRun saves them per window,
 * slosh2.c:2460-2475), env[nstorm][jmxb][imxb]; their rex frames come from
 * slg_batch_rex_frames(x, "HBMAX", ...) and their values from
 * slg_batch_get(x, s, "HBMAX", ...). */
int slg_batch_set_maxstat(slg_batch *x, int on);
int slg_batch_maxstat_reset(slg_batch *x);
int slg_batch_maxstat_env(slg_batch *x, short *env, char *err, int errlen);
/* The second window of the -envSave2 mode (hb_max2, slosh2.c:1264-1311): the
 * same maxima over a window of its own, which PerformRun restarts only when
 * an envSave2 save falls due (MaxStatInitGrid flag 3, slosh2.c:2514-2520) and
 * saves to the envDir2 files (slosh2.c:2477-2508, 2639-2668). */
int slg_batch_set_maxstat2(slg_batch *x, int on);
int slg_batch_maxstat2_reset(slg_batch *x);
int slg_batch_maxstat2_env(slg_batch *x, short *env, char *err, int errlen);
/* The file around the frames, laid out as the reference's rex writer does:
 * open = RexOpen (version 2) + RexSaveHeader (rex.c:150-180, 264-379); frame =
 * RexSaveStep's record (rex.c:387-447: lat, lon, wind speed, wind direction,
 * delp, size as the C driver holds them, and the model clock in seconds since
 * 1970); env = RexSaveEnv (rex.c:505-650, with Data_WriteTrk's copy of the
 * track file, rex.c:57-230, and RexSetMinMaxHt, rex.c:232-256). */
typedef struct slg_rex slg_rex;
int slg_rex_open(const char *path, const char *comment, const char *abbrev, int imxb, int jmxb, slg_rex **out);
int slg_rex_frame(slg_rex *r, const double track[6], double clock, const unsigned char *bits, int nbytes);
int slg_rex_env(slg_rex *r, const char *trk_path, const unsigned char *bits, int nbytes);
void slg_rex_close(slg_rex *r);
/* Rex point series without the frames.  slg_batch_rex_points gives the value
 * a rex frame of field ("HHB", "HMX" or "HBMAX", as slg_batch_rex_frames)
 * holds at each of npts squares xy[2p], xy[2p+1] (1-based I, J, at most
 * IMXB-1, JMXB-1): tenths of a foot, -320..700, 999 dry; out[ns][npts].
 * These are the values the reference's rexout reads back from a rex file
 * (sorc/rexout.cd/rexout.c:707-740).  slg_pnt_* write them as rexout's
 * style 0 text (rexout.c:631-639, 736-742), one row per frame with the
 * model clock's date. */
int slg_batch_rex_points(slg_batch *x, const char *field, int s0, int ns, int npts, const int *xy, short *out,
                         char *err, int errlen);
typedef struct slg_pnt slg_pnt;
int slg_pnt_open(const char *path, const char *abbrev, const char *storm, const char *rexfile, int npts,
                 const int *xy, slg_pnt **out);
int slg_pnt_row(slg_pnt *p, double clock, const short *vals);
void slg_pnt_close(slg_pnt *p);
/* Write an env file in EnvSave's layout (slosh2.c:1611-1671). */
int slg_write_env(const char *path, int imxb, int jmxb, const char *comment, const short *vals, float ht1,
                  float ht2);

/* ---- Serving: storm runs on several basins admitted as they arrive ----------
 * No counterpart in the reference, which runs one storm per process (slosh2.c
 * PerformRun) and queues whole runs in shell scripts (dev/a1.multiRun.sh).
 * A server holds one pool of run slots per registered basin, sized from free
 * HBM; runs wait in a FIFO until a slot of their basin is free, then step
 * together with the runs already there (continuous batching); every pump
 * retires finished runs (their envelope is computed on the device), admits
 * waiting runs and launches up to k further steps on every pool. */
typedef struct slg_server slg_server;
int slg_server_open(int precision, double hbm_fraction, slg_server **out, char *err, int errlen);
/* Register a basin with up to max_slots (<= SLG_MAX_STORMS) slots of max_steps steps and an arena
 * of tables_per_slot radial-table sets per slot; *basin_id identifies it. */
int slg_server_add_basin(slg_server *sv, const char *dta_path, const char *bnt_line, int max_slots,
                         int max_steps, int tables_per_slot, int *basin_id, char *err, int errlen);
int slg_server_slots(const slg_server *sv, int basin_id);
int slg_server_submit(slg_server *sv, int basin_id, const char *trk_path, int *run_id, char *err, int errlen);
/* Admission order of the queued runs: SLG_ORDER_FIFO (submission order, the
 * default) or SLG_ORDER_HARDEST (the most work first, steps x cells, as the
 * reference's job list is sorted for its worker queue: dev/b1.sortJobList.sh:
 * 20-73 with leader.c:48-71 handing idle workers the sorted list). */
#define SLG_ORDER_FIFO 0
#define SLG_ORDER_HARDEST 1
int slg_server_set_order(slg_server *sv, int order);
/* Returns the number of runs retired this round (>= 0) or an error code.  The
 * host waits for no stream: admissions and envelopes pass through pinned
 * staging behind events, and a pool is held to two pumps of queued steps. */
int slg_server_pump(slg_server *sv, int k, char *err, int errlen);
/* 0 queued, 1 running, 2 done (then env [jmxb][imxb] and latency are filled),
 * 3 failed (its track could not be read, it does not fit the pool's slots, or
 * its drying-cascade replay ran out of visits); slg_server_run_error gives the
 * message. */
int slg_server_result(const slg_server *sv, int run_id, short *env, double *latency_s);
int slg_server_run_error(const slg_server *sv, int run_id, char *err, int errlen);
/* Tile rows (8, 16 or 32) each basin's pool steps with, out[n]: every pool
 * times the three tiles on its first launch rounds (real steps; identical
 * fields) and keeps the fastest per slot-step.  Returns the number of pools. */
int slg_server_tiles(const slg_server *sv, int *out, int n);
/* [0] queued [1] running [2] done [3] slots [4] pool bytes [5] storm.cell.steps */
int slg_server_stats(const slg_server *sv, long long *out6);
void slg_server_close(slg_server *sv);

/* ---- tide coupling through the C driver (slosh2.c:813-1027) ----------------
 * The reference's driver adds the astronomical tide between steps: to the water
 * level of every wet square (COMMON HB) and as TIDESH at every BDRYHT square
 * (COMMON /BCPTS/, bdryht.f:85).  The Fortran-ABI library passes both on with
 * these calls; a batch can use them the same way.
 * boundary: the BDRYHT squares (ISH, JSH, 1-based; DEPSFC, depsfc.f:376-390), up
 * to cap of them; returns their number.  is_subgrid: ISSUBGRID (issubgrid.f:
 * 14-33).  put: overwrite a storm's HB (double [jmxb][imxb]) where its next
 * step reads it.  set_boundary_tide: TIDESH of a storm from its next step on,
 * n = the number of BDRYHT squares (0 until set).  static_height: STATICHT
 * (staticht.f:14-41) at (i,j) with the storm values of its next step. */
int slg_basin_boundary(const slg_basin *b, int *ish, int *jsh, int cap);
int slg_basin_is_subgrid(const slg_basin *b, int i, int j);
int slg_batch_put(slg_batch *x, int storm, const char *name, const double *src, char *err, int errlen);
int slg_batch_set_boundary_tide(slg_batch *x, int storm, const double *tide, int n, char *err, int errlen);
int slg_batch_static_height(const slg_batch *x, int storm, int i, int j, double *ht);

/* The gridded astronomical tide of the -f_tide V2 modes inside the batch, with
 * no driver between steps: the tide coupling of RunInit and RunLoopStep
 * (slosh2.c:1789-1869, 1553-1582), SetTideUsableFlag (:717-744) and
 * CalcTideGrid2 (:813-1027), applied on the device after every step (and once
 * to the initial water at the upload).  Replaces the reference's
 * InitTideGrid / ReadHarmonicBody / ReadAdjDatum (tideutil.c:148-734) and
 * CalcTideGrid2 per run.
 * ft03: the yearly constituent file; bhc: the basin's harmonic constants
 * (HarmConstit v1, one record per square of the (IMXB-1) x (JMXB-1) grid);
 * adj: the MTL -> NAVD88 adjustment text, or NULL / a missing file (0).
 * mode: 2, 21, 22 or 24 (-f_tide V2, V2.1.x, V2.2.x, V2.4.x) and thresh the x
 * (ParseTide, usrparse.c:14-52; squares with ZB <= thresh take no tide in
 * modes 21, 22, 24).  Each run's model clock starts at its start date and
 * moves by its DELT each step; the tables are its year's.  Call once, before
 * slg_batch_upload.  hipGraphs are not used while the tide is on (the steps
 * launch one by one); storm parts on several streams are (each part sums its
 * own tide fields).  The hsf basins call the tide every 6 model minutes, as
 * RunLoopStep and PerformRun do there (slosh2.c:1565-1577, 2583-2586).  The
 * tide spin-up (-spinUp) is not provided.
 * datums_tide: slg_track_datums for such a run (InitWater_TideModeOverride
 * with f_tide = mode). */
int slg_batch_set_tide(slg_batch *x, const char *ft03, const char *bhc, const char *adj, int mode, int thresh,
                       char *err, int errlen);
int slg_track_datums_tide(const char *trk_path, int mode, float *ht1, float *ht2);

/* Write the host initialisation of one run, and its storm schedule, in the
 * tagged-record layout used for the reference's state dumps (for checking). */
int slg_export_init(const slg_batch *x, int storm, const char *path);
int slg_export_storm(const slg_batch *x, int storm, const char *rows_path, const char *tables_path);

#ifdef __cplusplus
}
#endif
#endif
