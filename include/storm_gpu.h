/*
 * storm_gpu.h -- C ABI of the MI355X (gfx950) Implicit Monte Carlo engine.
 *
 * Plain C: opaque handle, POD structs, pointers and sizes, int status codes
 * (0 = success) with the message of the last failure on this thread in
 * storm_gpu_last_error().  No torch, HIP or C++ types cross this boundary.
 * Status: 1 invalid argument or problem, 2 device or HIP failure, 4 a size
 * or capacity limit, 5 a physics error the reference also stops on
 * (RadiationIMC.hpp:771-817, Fleck factor) or the tally bound (a cell with
 * 2^28 segments in a step; the material state is then left as it was) --
 * the step stopped part-way, and the engine refuses further steps -- and 6 a
 * packet beyond 2^24 segments in one step.
 *
 * The entry points replace the serial manager and its IMC physics plugin of
 * the reference for one GPU:
 *   MonteCarloManagerSerial<T,Grid>      manager/MonteCarloManagerSerial.hpp:25-117
 *   RadiationIMC<...>                    radiation/RadiationIMC.hpp:332-556
 *   CombPopulationControl                population/CombPopulationControl.hpp:20-32
 *   SideTemperature / Vacuum / Rigid     boundary/SideTemperature.hpp:20-167,
 *                                        boundary/Vacuum.hpp:10-45, boundary/RigidBoundary.hpp:9-24
 * The C++ drop-in (include/storm_amd/MonteCarloManagerGPU.hpp) binds them.
 */
#ifndef STORM_GPU_H
#define STORM_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "storm_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define STORM_GPU_ABI_VERSION 8

typedef struct storm_gpu_engine storm_gpu_engine;

/* Arithmetic of the transport. */
enum {
  STORM_GPU_MIXED = 0, /* FP32 cell-local tracking, own deterministic math, 64-bit fixed-point tallies */
  STORM_GPU_FP64 = 1   /* plain FP64 tracking, device libm, FP64 atomic tallies                      */
};

/* Where the per-cell opacities, heat capacities and temperatures come from. */
enum {
  STORM_GPU_MATERIAL_DEVICE = 0, /* storm_material_desc evaluated on the GPU every step      */
  STORM_GPU_MATERIAL_HOST = 1    /* host-tabulated plugins via storm_gpu_set_cell_state()   */
};

/* A census packet as the reference's Particle carries it
 * (particle/Particle.hpp:67-85): global position, velocity = c * direction,
 * cell, weight (erg), frequency. */
typedef struct storm_packet_record {
  double x, y, z;
  double vx, vy, vz;
  double weight;
  double frequency;
  uint64_t cell;
  uint64_t id;
} storm_packet_record;

/* Counters of the last step (MonteCarloManagerSerial.hpp:74-83 getters and
 * RadiationIMC counters, radiation/RadiationIMC.hpp:444-446). */
typedef struct storm_gpu_counters {
  uint64_t initial_particles;  /* census entering the step (GetInitialParticleCount)  */
  uint64_t prestep_particles;  /* emission + boundary sources (GetPreStepParticleCount) */
  uint64_t start_particles;    /* GetStartParticleCount                                 */
  uint64_t end_particles;      /* census after population control (GetEndParticleCount) */
  uint64_t emitted_particles;
  uint64_t boundary_particles;
  uint64_t census_before_comb;
  uint64_t leaving_particles;  /* escapes through REMOVE faces                          */
  uint64_t segments;           /* sum of GetCellsStepsCounters()                        */
  uint64_t transport_passes;
  uint64_t transport_launches;
  uint64_t random_walk_steps;  /* getRandomWalkStepCount */
  uint64_t ddmc_steps;         /* getDDMCStepCount       */
  int32_t energy_quantum_exp;  /* Q = 2^eq erg of the last step */
  int32_t transport_block;     /* block size the transport ran with in the last step */
  double seconds_step;         /* host wall time of the last storm_gpu_step          */
  double seconds_transport;    /* GPU time of the transport kernels (HIP events)     */
  double seconds_transport_max_launch; /* longest single transport launch              */
  uint64_t max_packet_blocks;  /* most random-number blocks one packet used this step
                                  (about its segments): the longest history */
  int32_t transport_chunks;    /* chunks per wavefront of the last step's transport */
  int32_t transport_schedule;  /* its schedule (STORM_GPU_SCHEDULE_*)               */
} storm_gpu_counters;

/* Exact energy ledger of the last step, in units of its quantum 2^eq erg
 * (mixed engine; the FP64 engine reports the same terms as rounded doubles).
 * Each term is hi * 2^64 + lo. */
typedef struct storm_gpu_u128 {
  uint64_t lo, hi;
} storm_gpu_u128;

typedef struct storm_gpu_ledger {
  storm_gpu_u128 census_in;   /* census weight entering transport          */
  storm_gpu_u128 emitted;     /* material emission packets                 */
  storm_gpu_u128 source;      /* boundary source packets                   */
  storm_gpu_u128 deposited;   /* absorbed in the material during transport */
  storm_gpu_u128 escaped;     /* left through REMOVE faces                 */
  storm_gpu_u128 census_pre;  /* census weight before population control   */
  storm_gpu_u128 census_out;  /* census weight after population control    */
  storm_gpu_u128 comb_rem;    /* comb rounding remainder (to the material) */
  storm_gpu_u128 comb_dropped; /* stratified comb: weight of the bins it drops (left the problem,
                                  as in CombPopulationControl.hpp:348-349, 409-413) */
  int32_t eq;                 /* quantum exponent                          */
  int32_t eq_prev;
  double requant_rem;         /* census requantisation remainder, erg      */
  double clamp;               /* negative-energy guard, erg (RadiationIMC.hpp:771-817) */
  double material_before, material_after; /* sum of cell internal energies, erg */
} storm_gpu_ledger;

const char *storm_gpu_last_error(void);
int storm_gpu_abi_version(void);

/* Device info string (name, CUs, arch) for logs. */
int storm_gpu_device_info(char *buf, size_t len);

/* Build the engine for `prob` (copied).  Replaces the manager constructor
 * (MonteCarloManagerSerial.hpp:119-129) together with the RadiationIMC
 * constructor (radiation/RadiationIMC.hpp:562-640). */
int storm_gpu_create(const storm_problem *prob, int32_t precision, int32_t material_mode,
                     storm_gpu_engine **out);

/* ~MonteCarloManagerSerial (MonteCarloManagerSerial.hpp:185-194). */
int storm_gpu_destroy(storm_gpu_engine *eng);

/* One IMC step of length dt with the census kept in HBM:
 * MonteCarloManagerSerial::step (MonteCarloManagerSerial.hpp:385-469) with
 * RadiationIMC::preStep/step/postStep (radiation/RadiationIMC.hpp:1618-2029),
 * the boundary sources and population control. */
int storm_gpu_step(storm_gpu_engine *eng, double dt);

/* storm_gpu_step without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code that runs beside HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc; the
 * step itself synchronises with the host, so it is not captured).  The same
 * step, bit for bit, with every transient buffer of the step (packets in
 * flight, the census before population control, sort keys, prefix sums,
 * reduction partials, counters) carved out of a fixed 64 MiB device array of
 * the library (one per GPU, held until the process ends); it calls no
 * hipMalloc, hipFree or other allocating routine.  Limits, each a clean error
 * (status 4) that leaves the census and material state unchanged: the step's
 * transient buffers must fit the 64 MiB (about 2e5 packets per step in the
 * mixed engine's refill schedule, which stages the census, 2.5e5 in its pass
 * schedule, fewer in FP64 or multigroup), and the census must fit the
 * engine's census buffer, which only storm_gpu_step and
 * storm_gpu_import_census grow (at least doubling).  All engines on one GPU
 * share the array: calls that overlap on one GPU race; serialise them.
 * Rejects an engine with the debug copy of transport inputs switched on. */
int storm_gpu_step_noalloc(storm_gpu_engine *eng, double dt);

/* SideTemperature::SetTemperature (boundary/SideTemperature.hpp:48) and the
 * MarshakBoundary equivalent. */
int storm_gpu_set_source_temperature(storm_gpu_engine *eng, double T);

/* Host-tabulated material (STORM_GPU_MATERIAL_HOST): per-cell temperature,
 * Planck/absorption opacity, scattering opacity, dT2cv and internal energy
 * for the next step; replaces the per-cell virtual calls of preStep
 * (radiation/RadiationIMC.hpp:1645-1648) and step (:1753).  Any pointer may
 * be NULL to keep the current values. */
int storm_gpu_set_cell_state(storm_gpu_engine *eng, const double *T, const double *sigma_a,
                             const double *sigma_s, const double *cv,
                             const double *internal_energy);

/* Per-cell results after the last step (any pointer may be NULL):
 * temperature (cells[i].temperature, RadiationIMC.hpp:848), internal energy
 * (extensives[i].internal_energy), time-averaged radiation energy density
 * (getEradTimeAvg, RadiationIMC.hpp:431, :1895), census radiation energy
 * (extensives[i].Erad, :1997), Fleck factor (getFactorFleck, :429) and
 * Planck opacity (getPlanckOpacities, :430). */
int storm_gpu_get_cells(storm_gpu_engine *eng, double *T, double *internal_energy,
                        double *erad_time_avg, double *census_erad, double *fleck,
                        double *planck_opacity);

/* Time-averaged radiation energy density of the last step per cell and
 * photon group, out[cell * G + g] (erg/cm^3): getEgTimeAvg
 * (RadiationIMC.hpp:433-434, tallied at :1166-1173, :1485-1492, :1802-1809,
 * normalised at :1896-1903).  Needs a multigroup problem created with
 * storm_group_desc.with_eg_time_avg (withEgTimeAvg); status 1 otherwise. */
int storm_gpu_get_group_erad(storm_gpu_engine *eng, double *out);

/* GetCellsStepsCounters (MonteCarloManagerSerial.hpp:74-75). */
int storm_gpu_get_segments(storm_gpu_engine *eng, uint64_t *per_cell);

/* GetBeginningParticleCount (MonteCarloManagerSerial.hpp:81-82). */
int storm_gpu_get_beginning_counts(storm_gpu_engine *eng, uint64_t *per_cell);

/* Counters of the last step. */
int storm_gpu_get_counters(storm_gpu_engine *eng, storm_gpu_counters *out);

/* Exact energy ledger of the last step.  In a batch (storm_problem
 * instances) the integer fields count the step's quantum in each instance's
 * own energy unit: the balances hold exactly, and the erg figures come from
 * storm_gpu_get_instance_ledger. */
int storm_gpu_get_ledger(storm_gpu_engine *eng, storm_gpu_ledger *out);

/* The ledger of the last step restricted to instance k of a batch (its cells
 * and source faces), with eq and eq_prev the instance's quantum exponents in
 * erg.  The reference steps one problem per manager, whose ledger this is
 * (MonteCarloManagerSerial.hpp:385-469). */
int storm_gpu_get_instance_ledger(storm_gpu_engine *eng, uint32_t k, storm_gpu_ledger *out);

/* Census exchange with the host -- the std::vector<Particle> moved into and
 * returned from MonteCarloManagerSerial::step (:44, :197-248, :464-468). */
int storm_gpu_census_size(storm_gpu_engine *eng, uint64_t *n);
int storm_gpu_export_census(storm_gpu_engine *eng, storm_packet_record *out, uint64_t capacity,
                            uint64_t *n_written);
int storm_gpu_import_census(storm_gpu_engine *eng, const storm_packet_record *in, uint64_t n);

/* New packets made on the host for the next step: the packets a user
 * boundary's generateNewBoundaryParticles returns, which RadiationIMC::preStep
 * adds to the step's sources (RadiationIMC.hpp:1683-1692), each with its own
 * timeLeft (time_left[i], s; MarshakBoundary draws dt * U,
 * examples/marshak_wave/MarshakBoundary.hpp:87-130).  Weights in erg; the
 * next storm_gpu_step puts them on its energy quantum, counts them as
 * boundary source in the ledger and transports them with the step's other
 * packets.  Calls accumulate until that step. */
int storm_gpu_queue_source_packets(storm_gpu_engine *eng, const storm_packet_record *in, const double *time_left,
                                   uint64_t n);

/* Launch parameters of the transport: block size (64, 128, 256, 512; 0:
 * tuned on line, below), segments per packet and launch, and the schedule
 * (-1: keep the engine's).  PASS runs each lane's packet to its end (or the
 * segment budget) before the wavefront takes the next 64 of the block's
 * range; REFILL advances the wavefront's packets one segment at a time and
 * hands lanes whose packet ended the next packets of the block's range, four
 * or more lanes at a time.  A block's range is block size x chunks per
 * wavefront (storm_gpu_set_chunks_per_wave; each schedule keeps its own,
 * 2 for PASS and 8 for REFILL by default, 16 for REFILL in the FP64 engine;
 * in the mixed engine on polyhedral meshes 4 for REFILL, and 1 for PASS on
 * multigroup problems; 16 for REFILL, in 512-thread blocks, on Cartesian
 * problems with DDMC; blocks are otherwise 256 threads; a launch halves
 * it while it would have fewer blocks than the device holds at once).  The engine starts in REFILL, except the mixed engine
 * on multigroup polyhedral problems (PASS), and watches: after 16 steps one
 * step runs in the other schedule, and the engine switches when that one
 * moved more transport segments per second than the mean of the steps
 * either side of it, by more than 3%; the next probe follows 16 steps later,
 * or twice the last interval (at most 128) when the probe lost by more than
 * 10%.  A block size other than 0 fixes the launch and stops the watch.  A packet's history does not depend on the schedule, the block
 * size or the chunks per wavefront, only on segments_per_pass: the mixed
 * engine's results are bit-identical across them, the FP64 engine's tallies
 * differ only in the order of their floating-point sums.  With block size 0
 * the engine tunes the launch on line, one parameter at a time: the schedule
 * (each with its chunks), the block size (128, 256, 512), then the chunks per
 * wavefront (1, 2, 4, 8, 16); each candidate runs one step forwards and one
 * backwards (an order that cancels a steady trend of the workload) and the
 * one with the highest mean rate of transport segments per second is kept
 * (20 steps; storm_gpu_get_launch shows the choice), after which the watch
 * runs.  The results are those of any fixed launch.  No counterpart in the
 * reference, whose manager has no launch to tune. */
#define STORM_GPU_SCHEDULE_PASS 0
#define STORM_GPU_SCHEDULE_REFILL 1
int storm_gpu_set_tuning(storm_gpu_engine *eng, int32_t block_size, int32_t segments_per_pass, int32_t schedule);

/* 64-packet chunks of the cell-ordered input per wavefront that a transport
 * block owns and hands out in the current schedule (1, 2, 4, 8 or 16): the
 * block's share of the input and of its LDS tally table.  No counterpart in
 * the reference. */
int storm_gpu_set_chunks_per_wave(storm_gpu_engine *eng, int32_t chunks);

/* The transport launch in use: block size, chunks per wavefront, schedule,
 * and whether on-line tuning is still running (1) or done (0).  No
 * counterpart in the reference. */
int storm_gpu_get_launch(storm_gpu_engine *eng, int32_t *block_size, int32_t *chunks_per_wave, int32_t *schedule,
                         int32_t *tuning);

/* Keep a copy of the next step's transport inputs (initial packets, cell
 * opacities, step constants) and its integer tallies, for an exact CPU replay
 * of the packet random streams.  Sizes first with NULL buffers.  cell_xs
 * receives two numbers per cell in the tracking type (f sigma_a and the
 * effective-scatter opacity).  After a step
 * fails because a packet exceeded the segment limit (error 6), the kept
 * packets are the (first 4096) packets still in flight instead. */
int storm_gpu_debug_keep_transport_inputs(storm_gpu_engine *eng, int32_t enable);
int storm_gpu_debug_transport_inputs(storm_gpu_engine *eng, uint64_t *n_packets, void *packets,
                                     void *cell_xs, int64_t *consts4);
/* Integer tallies in quanta of the step, two words per cell (value =
 * hi * 2^36 + lo): absorbed, rouletted, path integral (cells without
 * absorption only), escaped; and the segment counts. */
int storm_gpu_debug_tallies(storm_gpu_engine *eng, uint64_t *dep, uint64_t *rdep, uint64_t *pint, uint64_t *esc,
                            uint32_t *segs);
/* Per-cell random-walk / DDMC data of the last step (bytes written to
 * `out`, which may be NULL to query the size). */
int storm_gpu_debug_accel(storm_gpu_engine *eng, void *out, uint64_t *bytes);
/* Multigroup data of the last step in the engine's tracking type (float for
 * STORM_GPU_MIXED): group count, energy exponent, bounds in keV (groups + 1),
 * per-cell emission CDFs (cells * (groups + 1)) and scattering opacities.
 * Pointers may be NULL; nothing is written but *groups when gray. */
int storm_gpu_debug_spectrum(storm_gpu_engine *eng, int32_t *groups, double *nu_pow, void *bounds_kev, void *cdf,
                             void *sig_s);

/* Measurement aid with no counterpart in the reference: one plain copy of
 * `bytes` (a multiple of 16) between two fresh device buffers by a kernel
 * that reads and writes each byte once (k_copy_calibration), so that a
 * profiler's HBM byte counters can be scaled against known traffic.  *ms
 * (may be NULL) receives the kernel time. */
int storm_gpu_calibration_copy(uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* STORM_GPU_H */
