// remc_facade.h -- C++ drop-in for the reference's host-side API of the REMC
// path: AminoAcids, Molecule, Replica, the simulation-config file, the REMC
// driver and its output files.  Compute goes through the C ABI (remc.h).
//
// Mirrors (paths relative to "simulation engine/src/"):
//   AminoAcids           AminoAcids.cpp:25-195
//   Molecule             Molecule.h:77-95, Molecule.cpp:88-346,430-830
//   Replica              Replica.h, Replica.cpp:145-530,603-745,1717-1780
//   argdata/getArgs      definitions.h:203-225, main.cpp:193-381
//   loadArgsFromFile     main.cpp:383-632
//   REMCSimulation       main.cpp:1106-1567, initSamplingFiles main.cpp:634-713
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "remc.h"

namespace remc_facade {

constexpr float CROWDER_IDENTIFIER = -1.0f;   // definitions.h:92
constexpr float BOUND_ENERGY_VALUE = -1.1844f; // definitions.h:84
constexpr float MIN_TRANSLATION = 0.5f, MAX_TRANSLATION = 5.0f; // definitions.h:79-80
constexpr float MIN_ROTATION = 0.1f, MAX_ROTATION = 0.5f;       // definitions.h:82-83

struct Vector3f {
    float x = 0, y = 0, z = 0;
    Vector3f() = default;
    Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
};

struct Vector3double {
    double x = 0, y = 0, z = 0;
    Vector3double() = default;
    Vector3double(double a, double b, double c) : x(a), y(b), z(c) {}
    void normalizeInPlace(); // vector3f.cpp:36-43
};

struct Quaternion {
    double w = 1, x = 0, y = 0, z = 0;
    Quaternion() = default;
    Quaternion(double w_, double x_, double y_, double z_) : w(w_), x(x_), y(y_), z(z_) {}
    Quaternion normalize();                      // Quaternion.cpp:19-27
    Quaternion operator*(const Quaternion &q) const; // Quaternion.cpp:35-53
    Vector3f rotateVector(Vector3f v) const;     // Quaternion.cpp:84-101
};

struct AminoAcid {
    std::string name, sname;
    char symbol = ' ';
    float vanderWaalRadius = 0.f;
    float electrostaticCharge = 0.f;
};

class AminoAcids {
  public:
    AminoAcid data[REMC_NTYPES];
    float LJpotentials[REMC_NTYPES][REMC_NTYPES] = {};
    int count = 0;
    // The reference exits on a missing file; these return false instead.
    bool loadAminoAcidData(const char *filename);
    bool loadLJPotentialData(const char *filename);
    // -1 when unknown (the reference returns -1 too)
    int getAminoAcidIndex(const char *name) const;
    const AminoAcid &get(int i) const { return data[i]; }
};

struct Residue {
    int aminoAcidIndex = 0;
    float electrostaticCharge = 0.f;
    float vanderWaalRadius = 0.f;
    Vector3f position, relativePosition;
    char chainId = ' ';
    int resSeq = 0;
};

class Molecule {
  public:
    std::vector<Residue> Residues;
    Vector3f center, position;
    Quaternion rotation;
    float moleculeRoleIdentifier = 0.f;
    std::string filename;
    int index = 0;

    bool initFromPDB(const char *pdbfilename, const AminoAcids &aas);
    bool translate(Vector3f v);
    void setPosition(Vector3f v);
    bool rotateQ(Vector3double axis, double angle);
    void setRotation(Quaternion q);
    void setMoleculeRoleIdentifier(float role);
    // Molecule.cpp:88-102: residues, centre, position and rotation copied from
    // m (the state before a move), and copied back from m to undo the move
    void saveBeforeStateChange(const Molecule *m);
    void undoStateChange(const Molecule *m);
    bool isCrowder() const { return moleculeRoleIdentifier == CROWDER_IDENTIFIER; }
    size_t residueCount() const { return Residues.size(); }
};

// parameters of a run (definitions.h:203-225)
struct argdata {
    int threads = 1, streams = 1, gpus = 1;
    int MCsteps = 5000, REsteps = 10, replicas = 20;
    float bound = 150.0f;
    int sampleFrequency = 5000, sampleStartsAfter = 5000;
    std::string prependageString = "output";
    std::string file;
    bool inputFile = false;
    int nonCrowders = 0;
    float temperatureMin = 250.0f, temperatureMax = 500.0f;
    int pid = 0;
    uint64_t seed = 0;
    bool seedGiven = false;
    std::string logfile;
    std::string outputDir = "output";
    std::string dataDir;
    int device = 0;
    int blockDim = 0; // MC workgroup size; 0 = autotuned
    bool dryRun = false;
    bool help = false;
    bool check = false;            // -c: plain against batched kernel on the example conformations
    std::string checkDir = "data"; // holds conf1/1a.pdb, conf1/1b.pdb, ... conf10/1b.pdb
    std::string savePdb, finalPdb; // initial / final (lowest temperature) conformation
    int coop = 0;                  // MC workgroups per replica; 0 = autotuned
    int checkpointFrequency = 0;   // MC steps between checkpoints; 0 = none
    std::string checkpointDir = "checkpoints";
    bool resume = false;
    std::string resumeFile;
};

// Statistics that follow a temperature rung: they swap with the temperature
// on an exchange (Replica.h:119-131, Replica.cpp:145-159).
struct SlotStats {
    long totalAccept = 0, totalAcceptReject = 0, boundSamples = 0, samplesSinceLastExchange = 0,
         totalBoundSamples = 0, totalSamples = 0;
};

// Everything a run needs to continue bit for bit: the reference's checkpoint
// content (saveSimulation, main.cpp:89-117) plus the exact state (bead
// coordinates, tracked energies, RNG step counters, ladder, exchange state).
struct Checkpoint {
    int roundsDone = 0, totalExchanges = 0, totalTests = 0;
    uint64_t seed = 0;
    float box = 0.f;
    std::vector<int> ladder; // slot -> replica
    uint64_t exRound = 0, exSwaps = 0;
    int exOffset = 0;
    std::vector<double> energy;
    std::vector<remc_counters> counters, seen; // device counters; values at the last report
    std::vector<SlotStats> slots;
    int nBeads = 0, nBodies = 0;
    std::vector<float> x, y, z, rx, ry, rz; // replica-major, nBeads each
    std::vector<float> cx, cy, cz;          // replica-major, nBodies each
    std::vector<double> quat;               // replica-major, 4 nBodies each
};

// Checkpoint file: the reference's layout (main.cpp:89-117) followed by the
// exact section, floating-point values in hexadecimal so they read back
// unchanged.  loadSimulation fills the run parameters and the state.
bool saveSimulation(const std::string &path, const argdata &p, const Checkpoint &ck);
bool loadSimulation(const std::string &path, argdata *p, Checkpoint *ck);

class Replica {
  public:
    // The replica's own device context (ReplicaDataToDevice ... FreeDevice).
    // A copy of a Replica does not share it: the copy starts without one.
    struct DeviceMirror {
        remc_ctx *ctx = nullptr;
        int sys = -1;
        remc_counters seen{};  // counters already folded into accept/acceptA/reject
        DeviceMirror() = default;
        DeviceMirror(const DeviceMirror &) {}
        DeviceMirror &operator=(const DeviceMirror &o)
        {
            if (this != &o) reset();
            return *this;
        }
        ~DeviceMirror() { reset(); }
        void reset();
    };

    std::vector<Molecule> molecules;
    AminoAcids aminoAcids;
    float temperature = 300.0f;
    int label = -1;
    double potential = 0.0;
    float boundingValue = 150.0f;
    float translateStep = 0.5f;
    double rotateStep = 0.2;
    int nonCrowderCount = 0;
    // statistics (Replica.h:119-131)
    long accept = 0, acceptA = 0, reject = 0, totalAccept = 0, totalAcceptReject = 0;
    long boundSamples = 0, samplesSinceLastExchange = 0, totalBoundSamples = 0, totalSamples = 0;
    int blockSize = 0;      // tile of the plain kernel (0: 32 below 1024 beads, else 64)
    uint64_t rngSeed = 0;   // key of the replica's counter-based RNG streams (initRNGs)
    DeviceMirror mirror;

    void setAminoAcidData(const AminoAcids &a) { aminoAcids = a; }
    int loadMolecule(const char *pdbfilename);
    int loadMolecule(const char *pdbfilename, Vector3f position, Vector3double rotationAxis, double rotationAmount);
    void setBoundingValue(float v) { boundingValue = v; }
    void setTemperature(float t) { temperature = t; }
    void setTranslateStep(float t) { translateStep = t; }
    void setRotateStep(double r) { rotateStep = r; }
    void setLabel(int l) { label = l; }
    int residueCount() const;
    long countpairs() const;
    void exchangeReplicas(Replica &r); // Replica.cpp:145-159
    // Full energy on the GPU (Replica::E / EonDevice); a single-replica context.
    double E(int device = 0);

    // Replica.cpp:161-216: label, temperature, amino acids, box, molecules,
    // potential and block size of r (not the device context)
    void copy(const Replica &r);
    // Replica.cpp:265-290.  The reference seeds seven GSL streams from the
    // clock; here one counter-based (Philox) stream family keyed by seed and
    // label, taken by the device context ReplicaDataToDevice creates.
    void initRNGs(uint64_t seed = 0);
    // Replica.cpp:1108-1121: the plain kernel's tile (32, 64, 128 or 256)
    void setBlockSize(int b) { blockSize = b; }
    // Replica.cpp:1308-1417: this replica's system and state on a device
    // context of its own; returns 0, or -1 with the library's message printed
    int ReplicaDataToDevice(int device = 0);
    // Replica.cpp:1530-1576: the host conformation (after host-side moves of
    // molecule moleculeId, or of any molecule) to the device mirror
    int MoleculeDataToDevice(int moleculeId);
    // Replica.cpp:1618-1655
    void FreeDevice();
    // Replica.cpp:1601-1616: the reference-shaped tiled kernel on the device
    // mirror of the host conformation (created on first use)
    double EonDevice();
    // Replica.cpp:368-478: steps MC steps on the device mirror (created on
    // first use) from the host conformation; the conformation, potential and
    // the accept / acceptA / reject counters come back.  Returns 0 or -1.
    int MCSearch(int steps);
    // Replica.cpp:1717-1780: counts the sample; E_complex = potential plus the
    // energy between non-crowder molecules (on the device mirror, accumulated
    // in float as the reference does); if E_complex < boundEnergyThreshold
    // (or recordAllSamples) appends the boundconformations record.  Returns
    // whether the sample is bound.
    bool sample(FILE *boundConformations, int current_step, float boundEnergyThreshold,
                bool recordAllSamples = false);
    void saveAsSinglePDB(const char *filename) const; // Replica.cpp:495-530
    // one MODEL of a multi-model trajectory ("post processor/postprocess/main.cpp:252-291")
    void writeToTrajectory(FILE *file, int frame, double drms) const;

    // system packing for the C ABI
    struct Packed {
        std::vector<int> body_start, type;
        std::vector<float> charge, sigma, x, y, z, rx, ry, rz, cx, cy, cz;
        std::vector<unsigned char> crowder;
        std::vector<double> quat;
        remc_system_desc desc();
    };
    Packed pack() const;
};

// Distance RMS between two conformations of the same molecules: the mean over
// inter-molecule residue pairs of | |r_i - r_j|_a - |r_i - r_j|_b |; with
// reverseOrder, the smaller of that and the value with b's two molecules
// swapped (two equal-sized molecules) ("post processor/pdbgensrc/drms.cpp:7-58").
double DRMS(const Replica &sim, const Replica &exp, bool reverseOrder);

// force field of an AminoAcids table with the reference constants (definitions.h:18-35)
remc_forcefield makeForceField(const AminoAcids &aas);

// CLI (main.cpp:193-381, the options of this path) and config file
// (main.cpp:383-632).  Return false on a fatal error (message printed).
bool getArgs(argdata *d, int argc, char **argv);
bool loadArgsFromFile(argdata *d, Replica *initial);

// -c (main.cpp:1652-1707): the ten example conformations <checkDir>/conf<i>/1a.pdb
// + 1b.pdb in a 118.4 A box, each summed by the reference-shaped plain kernel
// and by the batched full-energy kernel; prints one "PLAIN,BATCHED,err:" line
// per conformation and returns 0, or 1 when a file or device call fails.
int checkConformations(const AminoAcids &aas, const argdata &p);

// The start-up consistency check (main.cpp:1767-1789): E of `initial` by the
// batched kernel (returned, NAN on failure) against the plain kernel, with
// the absolute and relative error and the reference's warning above 1 %.
double checkedInitialEnergy(Replica *initial, int device);

// The REMC driver with the reference's sampling and output files
// (output/<prefix>_<pid>_{fractionBound,boundconformations,acceptance_ratios,
// exchange_freq}); returns 0 on success.  final_replicas (may be null)
// receives every replica's end state and statistics; resume (may be null)
// continues a run from its checkpoint.
int REMCSimulation(Replica *initial, argdata *params, std::vector<Replica> *final_replicas = nullptr,
                   const Checkpoint *resume = nullptr);

} // namespace remc_facade
