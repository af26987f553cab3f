// Binary checkpoint of a context's time-stepping state (replaces the Tecplot
// ASCII restart of RunAgain, Solver_Mod.f90:164-195 with
// PrintResult.f90:192-270, which keeps 14 decimals and so does not resume
// bit for bit).
//
// The state is everything a time step reads from the step before it: the P
// interface (vof, phi, nx, ny), the U/V interface (vof, phi, nx, ny), the
// velocity and pressure, the Adams-Bashforth history Flux_n1, the
// predictor's Dp of the U and V cells (a cell that is not active keeps the
// last value written to it, and the next Poisson assembly reads Dp of every
// face between wet cells, ProjectionP_Mod.f90:418-449), the Poisson
// solution in the solver's layout (the start of a previous-pressure solve),
// the last momentum solutions (the start of the next momentum solves) and
// the particle set.  The geometry is not part of it: a restart sets the
// same geometry up first, as the reference recomputes it.  Layout: a fixed
// header, then each piece's bytes in the order of the header's list.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr char MAGIC[8] = {'S', 'G', 'S', 'T', 'A', 'T', 'E', '1'};
constexpr int32_t VERSION = 2;
constexpr int STATE_FIELDS[] = {SG_F_VOF,  SG_F_PHI,  SG_F_NX,  SG_F_NY,  SG_F_UVOF,    SG_F_VVOF,
                                SG_F_UPHI, SG_F_UNX,  SG_F_UNY, SG_F_VPHI, SG_F_VNX,    SG_F_VNY,
                                SG_F_U,    SG_F_V,    SG_F_P,   SG_F_FLUXN1U, SG_F_FLUXN1V,
                                SG_F_DPU,  SG_F_DPV};
constexpr int NSTATE = (int)(sizeof(STATE_FIELDS) / sizeof(STATE_FIELDS[0]));

struct Header {
    char magic[8];
    int32_t version, ni, nj, nfields;
    int32_t fields[32];
    int64_t solution_bytes;  // Poisson solution, solver layout
    int64_t momentum_bytes;  // each of the two momentum solutions
    int64_t particles;       // particle count (x, y, u, v, d of each follow)
};
static_assert(NSTATE <= 32, "header field list");
static_assert(sizeof(Header) == 176 && offsetof(Header, particles) == 168, "header layout is part of the format");

struct Piece {
    void* dev;
    size_t bytes;
};

size_t solution_bytes(const Ctx* c) { return c->lev[0].d.alloc_elems() * sizeof(double); }
size_t momentum_bytes(const Ctx* c) { return (size_t)c->ni * c->nj * sizeof(double); }

std::vector<Piece> pieces(Ctx* c, int64_t np) {
    std::vector<Piece> out;
    for (int k = 0; k < NSTATE; ++k) out.push_back({c->field[STATE_FIELDS[k]], c->field_bytes[STATE_FIELDS[k]]});
    out.push_back({c->x, solution_bytes(c)});
    out.push_back({c->pred.xu, momentum_bytes(c)});
    out.push_back({c->pred.xv, momentum_bytes(c)});
    const size_t pb = (size_t)np * sizeof(double);
    if (np > 0) {
        ParticleState& S = c->part;
        for (double* p : {S.x, S.y, S.u, S.v, S.d}) out.push_back({p, pb});
    }
    return out;
}

size_t total_bytes(Ctx* c, int64_t np) {
    size_t b = sizeof(Header);
    for (const Piece& p : pieces(c, np)) b += p.bytes;
    return b;
}

}  // namespace

}  // namespace sg

using namespace sg;

extern "C" {

int sg_state_bytes(sg_ctx* ctx, size_t* bytes) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !bytes) {
        set_error("sg_state_bytes: null argument");
        return SG_ERR_ARG;
    }
    *bytes = total_bytes(c, c->part.n);
    return SG_OK;
}

int sg_state_save(sg_ctx* ctx, void* buf, size_t bytes) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !buf) {
        set_error("sg_state_save: null argument");
        return SG_ERR_ARG;
    }
    const int64_t np = c->part.n;
    if (bytes < total_bytes(c, np)) {
        set_error("sg_state_save: buffer smaller than sg_state_bytes");
        return SG_ERR_ARG;
    }
    Header h{};
    std::memcpy(h.magic, MAGIC, sizeof(MAGIC));
    h.version = VERSION;
    h.ni = c->ni;
    h.nj = c->nj;
    h.nfields = NSTATE;
    for (int k = 0; k < NSTATE; ++k) h.fields[k] = STATE_FIELDS[k];
    h.solution_bytes = (int64_t)solution_bytes(c);
    h.momentum_bytes = (int64_t)momentum_bytes(c);
    h.particles = np;
    char* out = static_cast<char*>(buf);
    std::memcpy(out, &h, sizeof(h));
    size_t off = sizeof(h);
    for (const Piece& p : pieces(c, np)) {
        SG_HIP(hipMemcpyAsync(out + off, p.dev, p.bytes, hipMemcpyDeviceToHost, c->stream));
        off += p.bytes;
    }
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_state_load(sg_ctx* ctx, const void* buf, size_t bytes) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !buf || bytes < sizeof(Header)) {
        set_error("sg_state_load: need a buffer holding a saved state");
        return SG_ERR_ARG;
    }
    Header h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, MAGIC, sizeof(MAGIC)) != 0 || h.version != VERSION) {
        set_error("sg_state_load: not a state of this library version");
        return SG_ERR_ARG;
    }
    bool same = h.ni == c->ni && h.nj == c->nj && h.nfields == NSTATE &&
                h.solution_bytes == (int64_t)solution_bytes(c) && h.momentum_bytes == (int64_t)momentum_bytes(c) &&
                h.particles >= 0 && h.particles <= INT32_MAX;
    for (int k = 0; same && k < NSTATE; ++k) same = h.fields[k] == STATE_FIELDS[k];
    if (!same) {
        set_error("sg_state_load: the state was saved from a context of another grid");
        return SG_ERR_ARG;
    }
    const int64_t np = h.particles;
    if (bytes < total_bytes(c, np)) {
        set_error("sg_state_load: buffer shorter than the state it describes");
        return SG_ERR_ARG;
    }
    int rc = particles_reserve(c, (int)np);
    if (rc) return rc;
    c->part.n = (int)np;
    const char* in = static_cast<const char*>(buf);
    size_t off = sizeof(h);
    for (const Piece& p : pieces(c, np)) {
        SG_HIP(hipMemcpyAsync(p.dev, in + off, p.bytes, hipMemcpyHostToDevice, c->stream));
        off += p.bytes;
    }
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

}  // extern "C"
