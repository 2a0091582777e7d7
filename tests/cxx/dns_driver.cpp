// A compiled C++ caller of the cfgpu boundary: channelflow's DNS call pattern
// (tests/timeIntegrationTest.cpp:98-118, programs/simulateflow.cpp:131-142)
// written against include/cfgpu.hpp, with minimal stand-ins of channelflow's
// FlowField and DNSFlags that expose the members the facade uses.
//
//   dns_driver probe                          load the library, print the device count
//   dns_driver dns   Nx Ny Nz Lx Lz N in out  plane Couette, Re 400, skew-symmetric,
//                                             bulk velocity, SBDF3 from SMRK2, dt 0.02:
//                                             N steps as advance() calls of 5
//   dns_driver batch Nx Ny Nz Lx Lz N in out  the same integration as member 0 of a
//                                             DNSBatch beside a second member (u scaled by 1/2)
// in/out: the 3-component spectral velocity, channelflow's serial layout, raw doubles.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cfgpu.hpp"

namespace ref {  // stand-ins with channelflow's member names (flowfield.h, dnsflags.h)

enum BaseFlow { ZeroBase, LinearBase, ParabolicBase, LaminarBase, SuctionBase, ArbitraryBase };
enum MeanConstraint { PressureGradient, BulkVelocity };
enum TimeStepMethod { CNFE1, CNAB2, CNRK2, SMRK2, SBDF1, SBDF2, SBDF3, SBDF4 };
enum NonlinearMethod { Rotational, Convection, Divergence, SkewSymmetric, Alternating, Alternating_ };
enum Dealiasing { NoDealiasing, DealiasXZ, DealiasY, DealiasXYZ };

struct DNSFlags {
    BaseFlow baseflow = LaminarBase;
    MeanConstraint constraint = PressureGradient;
    TimeStepMethod timestepping = SBDF3;
    TimeStepMethod initstepping = SMRK2;
    NonlinearMethod nonlinearity = Rotational;
    Dealiasing dealiasing = DealiasXZ;
    bool taucorrection = true;
    double nu = 0.0025, Vsuck = 0, rotation = 0, dPdx = 0, dPdz = 0, Ubulk = 0, Wbulk = 0;
    double ulowerwall = 0, uupperwall = 0, wlowerwall = 0, wupperwall = 0, t0 = 0, dt = 0.01;
};

class FlowField {  // spectral state only
   public:
    FlowField(int Nx, int Ny, int Nz, int Nd, double Lx, double Lz, double a, double b)
        : Nx_(Nx), Ny_(Ny), Nz_(Nz), Nd_(Nd), Lx_(Lx), Lz_(Lz), a_(a), b_(b),
          d_((size_t)Nx * Ny * (Nz / 2 + 1) * Nd) {}
    int Nx() const { return Nx_; }
    int Ny() const { return Ny_; }
    int Nz() const { return Nz_; }
    int Nd() const { return Nd_; }
    double Lx() const { return Lx_; }
    double Lz() const { return Lz_; }
    double a() const { return a_; }
    double b() const { return b_; }
    void makeSpectral() {}
    std::complex<double>& cmplx(int mx, int my, int mz, int i) {
        return d_[mz + (size_t)(Nz_ / 2 + 1) * (mx + (size_t)Nx_ * (my + (size_t)Ny_ * i))];
    }
    size_t size() const { return d_.size(); }
    std::complex<double>* data() { return d_.data(); }

   private:
    int Nx_, Ny_, Nz_, Nd_;
    double Lx_, Lz_, a_, b_;
    std::vector<std::complex<double>> d_;
};

}  // namespace ref

static void read_field(const char* path, ref::FlowField& u) {
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(u.data()), (std::streamsize)(u.size() * sizeof(std::complex<double>)));
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
}

static void write_field(const char* path, ref::FlowField& u) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(u.data()), (std::streamsize)(u.size() * sizeof(std::complex<double>)));
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
}

int main(int argc, char** argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "probe") {
            int n = -1;
            const int rc = cfg_device_count(&n);
            std::printf("cfgpu loaded: cfg_device_count rc=%d n=%d\n", rc, rc == CFG_OK ? n : 0);
            return 0;
        }
        if ((mode != "dns" && mode != "batch") || argc != 10) {
            std::fprintf(stderr, "usage: %s probe | {dns|batch} Nx Ny Nz Lx Lz N in out\n", argv[0]);
            return 2;
        }
        const int Nx = std::atoi(argv[2]), Ny = std::atoi(argv[3]), Nz = std::atoi(argv[4]);
        const double Lx = std::atof(argv[5]), Lz = std::atof(argv[6]);
        const int N = std::atoi(argv[7]);
        ref::DNSFlags flags;  // plane Couette at Re 400 with the flags of the ensemble members
        flags.baseflow = ref::LaminarBase;
        flags.ulowerwall = -1.0;
        flags.uupperwall = 1.0;
        flags.constraint = ref::BulkVelocity;
        flags.Ubulk = 0.0;
        flags.nu = 1.0 / 400;
        flags.dt = 0.02;
        flags.timestepping = ref::SBDF3;
        flags.initstepping = ref::SMRK2;
        flags.nonlinearity = ref::SkewSymmetric;
        flags.dealiasing = ref::DealiasXZ;

        ref::FlowField u(Nx, Ny, Nz, 3, Lx, Lz, -1.0, 1.0), q(Nx, Ny, Nz, 1, Lx, Lz, -1.0, 1.0);
        read_field(argv[8], u);
        std::vector<ref::FlowField> fields = {u, q};
        if (mode == "dns") {
            cfgpu::DNS<ref::FlowField, ref::DNSFlags> dns(fields, flags);
            for (int done = 0; done < N; done += 5) {  // advance in blocks, as a saving loop does
                dns.advance(fields, std::min(5, N - done));
                std::printf("t = %.3f  CFL = %.6f  dPdx = %.6e  Ubulk = %.3e\n", dns.time(), dns.CFL(fields[0]),
                            dns.dPdx(), dns.Ubulk());
            }
            std::printf("Ubase[0..1] = %.6f %.6f\n", dns.Ubase()[0], dns.Ubase()[1]);
            write_field(argv[9], fields[0]);
        } else {
            ref::FlowField u2 = u;
            for (size_t k = 0; k < u2.size(); ++k) u2.data()[k] *= 0.5;
            std::vector<ref::FlowField> other = {u2, q};
            cfgpu::DNSBatch<ref::FlowField, ref::DNSFlags> batch;
            const int id0 = batch.add(fields, flags);
            const int id1 = batch.add(other, flags);
            batch.advance(N);
            batch.fetch(id0, fields);
            std::printf("members %d %d at t = %.3f, %.3f\n", id0, id1, batch.time(id0), batch.time(id1));
            write_field(argv[9], fields[0]);
        }
        return 0;
    } catch (const cfgpu::Error& e) {
        std::fprintf(stderr, "cfgpu error %d: %s\n", e.code(), e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
