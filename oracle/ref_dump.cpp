// Fixture dump driver for the reference GANDALF build (test infrastructure only).
//
// Links against the reference's own objects (oracle/Makefile.ref) and drives them through
// the public SimulationBase API exactly as src/Common/gandalf.cpp:40-140 does:
// Parameters::ReadParamsFile -> SimulationBase::SimulationFactory -> SetupSimulation ->
// MainLoop.  After setup and after each requested step it writes every per-particle field
// the GPU path must reproduce, at full precision, in a flat binary format:
//
//   int32 N, int32 nfields, then nfields arrays of N doubles, in FIELD_NAMES order.
//
// usage: ref_dump <paramfile> <nsteps> <outprefix> [dump_every]
//   writes <outprefix>_s<step>.bin (step 0 = after setup; gas), <outprefix>_stars_s<step>.bin
//   (stars: r v a adot m h gpot dt) and <outprefix>_diag.txt, and the
//   reference's own column, Seren unformatted, formatted and lite snapshots of each dumped
//   step after setup
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "Parameters.h"
#include "Simulation.h"
#include "Exception.h"
#include "Particle.h"
#include "Nbody.h"
#include "NbodyParticle.h"

static const char* FIELD_NAMES[] = {
  "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "atx", "aty", "atz",
  "m", "h", "rho", "u", "dudt", "div_v", "gpot", "invomega", "zeta", "sound",
  "pressure", "hfactor", "iorig", "levelneib", "dt", "level", "nstep", "nlast", "tlast"
};
static const int NFIELDS = sizeof(FIELD_NAMES)/sizeof(FIELD_NAMES[0]);

static void dump(Simulation<3>* sim, const std::string& fname)
{
  Hydrodynamics<3>* hydro = sim->hydro;
  const int N = hydro->Nhydro;
  std::vector<double> buf((size_t) N*NFIELDS);
  for (int i=0; i<N; i++) {
    GradhSphParticle<3>& p = static_cast<GradhSphParticle<3>&>(hydro->GetParticlePointer(i));
    double* f = &buf[0];
    int k = 0;
    for (int d=0; d<3; d++) f[(size_t)(k++)*N + i] = p.r[d];
    for (int d=0; d<3; d++) f[(size_t)(k++)*N + i] = p.v[d];
    for (int d=0; d<3; d++) f[(size_t)(k++)*N + i] = p.a[d];
    for (int d=0; d<3; d++) f[(size_t)(k++)*N + i] = p.atree[d];
    f[(size_t)(k++)*N + i] = p.m;
    f[(size_t)(k++)*N + i] = p.h;
    f[(size_t)(k++)*N + i] = p.rho;
    f[(size_t)(k++)*N + i] = p.u;
    f[(size_t)(k++)*N + i] = p.dudt;
    f[(size_t)(k++)*N + i] = p.div_v;
    f[(size_t)(k++)*N + i] = p.gpot;
    f[(size_t)(k++)*N + i] = p.invomega;
    f[(size_t)(k++)*N + i] = p.zeta;
    f[(size_t)(k++)*N + i] = p.sound;
    f[(size_t)(k++)*N + i] = p.pressure;
    f[(size_t)(k++)*N + i] = p.hfactor;
    f[(size_t)(k++)*N + i] = p.iorig;
    f[(size_t)(k++)*N + i] = p.levelneib;
    f[(size_t)(k++)*N + i] = p.dt;
    f[(size_t)(k++)*N + i] = p.level;
    f[(size_t)(k++)*N + i] = p.nstep;
    f[(size_t)(k++)*N + i] = p.nlast;
    f[(size_t)(k++)*N + i] = p.tlast;
  }
  FILE* fp = fopen(fname.c_str(), "wb");
  if (!fp) { perror(fname.c_str()); exit(2); }
  int hdr[2] = {N, NFIELDS};
  fwrite(hdr, sizeof(int), 2, fp);
  fwrite(&buf[0], sizeof(double), buf.size(), fp);
  fclose(fp);
}

// pure N-body runs (sim = nbody): the stars' state in the same flat format
static const char* STAR_FIELDS[] = {
  "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "adotx", "adoty", "adotz", "m", "h",
  "gpot", "dt"
};
static const int NSTARFIELDS = sizeof(STAR_FIELDS)/sizeof(STAR_FIELDS[0]);

static void dump_stars(Simulation<3>* sim, const std::string& fname)
{
  Nbody<3>* nbody = sim->nbody;
  const int N = nbody->Nstar;
  std::vector<double> buf((size_t) N*NSTARFIELDS);
  for (int i=0; i<N; i++) {
    const NbodyParticle<3>& p = nbody->stardata[i];
    int k = 0;
    for (int d=0; d<3; d++) buf[(size_t)(k++)*N + i] = p.r[d];
    for (int d=0; d<3; d++) buf[(size_t)(k++)*N + i] = p.v[d];
    for (int d=0; d<3; d++) buf[(size_t)(k++)*N + i] = p.a[d];
    for (int d=0; d<3; d++) buf[(size_t)(k++)*N + i] = p.adot[d];
    buf[(size_t)(k++)*N + i] = p.m;
    buf[(size_t)(k++)*N + i] = p.h;
    buf[(size_t)(k++)*N + i] = p.gpot;
    buf[(size_t)(k++)*N + i] = p.dt;
  }
  FILE* fp = fopen(fname.c_str(), "wb");
  if (!fp) { perror(fname.c_str()); exit(2); }
  int hdr[2] = {N, NSTARFIELDS};
  fwrite(hdr, sizeof(int), 2, fp);
  fwrite(&buf[0], sizeof(double), buf.size(), fp);
  fclose(fp);
}

static void dump_all(Simulation<3>* sim, const std::string& prefix, const std::string& tag)
{
  if (sim->hydro->Nhydro > 0) dump(sim, prefix + tag + ".bin");
  if (sim->nbody->Nstar > 0) dump_stars(sim, prefix + "_stars" + tag + ".bin");
}

static void diag_line(Simulation<3>* sim, FILE* fp)
{
  sim->CalculateDiagnostics();
  const Diagnostics<3>& d = sim->diag;
  // block-timestep structure: integer time, levels, and the particles that began a new step
  // at this integer time (nlast == n)
  Hydrodynamics<3>* hydro = sim->hydro;
  int nnew = 0;
  for (int i=0; i<hydro->Nhydro; i++) if (hydro->GetParticlePointer(i).nlast == sim->n) nnew++;
  fprintf(fp, "%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g"
          " %d %d %d %d %.17g %d\n",
          sim->Nsteps, (double) sim->t, (double) sim->timestep, d.Etot, d.ketot, d.utot, d.gpetot,
          d.mom[0], d.mom[1], d.mom[2], d.angmom[0], d.angmom[1], d.angmom[2], d.mtot,
          sim->n, sim->level_max, sim->level_step, sim->nresync, (double) sim->dt_max, nnew);
  fflush(fp);
}

int main(int argc, char** argv)
{
  if (argc < 4) {
    fprintf(stderr, "usage: %s <paramfile> <nsteps> <outprefix> [dump_every]\n", argv[0]);
    return 1;
  }
  const std::string paramfile = argv[1];
  const int nsteps = atoi(argv[2]);
  const std::string prefix = argv[3];
  const int every = argc > 4 ? atoi(argv[4]) : nsteps;

  ExceptionHandler::makeExceptionHandler(cplusplus);
  Parameters* params = new Parameters();
  params->ReadParamsFile(paramfile);
  SimulationBase* base = SimulationBase::SimulationFactory(params->intparams["ndim"],
                                                           params->stringparams["sim"], params);
  base->SetupSimulation();
  Simulation<3>* sim = static_cast<Simulation<3>*>(base);

  FILE* dfp = fopen((prefix + "_diag.txt").c_str(), "w");
  fprintf(dfp, "# Nsteps t timestep Etot ketot utot gpetot momx momy momz angx angy angz mtot"
          " n level_max level_step nresync dt_max nnew\n");
  diag_line(sim, dfp);
  dump_all(sim, prefix, "_s0");
  for (int s=1; s<=nsteps; s++) {
    sim->MainLoop();
    diag_line(sim, dfp);
    if (every > 0 && (s % every == 0 || s == nsteps)) {
      char name[64];
      snprintf(name, sizeof(name), "_s%d", s);
      dump_all(sim, prefix, name);
      snprintf(name, sizeof(name), "_s%d.col", s);
      sim->WriteColumnSnapshotFile(prefix + name);
      snprintf(name, sizeof(name), "_s%d.su", s);
      sim->WriteSerenUnformSnapshotFile(prefix + name);
      snprintf(name, sizeof(name), "_s%d.sf", s);
      sim->WriteSerenFormSnapshotFile(prefix + name);
      snprintf(name, sizeof(name), "_s%d.slite", s);
      sim->WriteSerenLiteSnapshotFile(prefix + name);
    }
  }
  fclose(dfp);
  return 0;
}
