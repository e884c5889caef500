// Harness around the reference's own hot-path closure (TEST INFRASTRUCTURE).
//
// Compiled by oracle/ref.mk against the reference sources where they lie
// (versions/v0.4/src of the reference tree) into oracle/_ref/; it never ships
// in the product.  It reads a case file written by tests/golden/make_golden.py,
// compiles the reference NativeCartesianMgPlan on MPI_COMM_SELF, and writes:
// the level plan, every level's diagonal, one preconditioner application with
// all level-0 workspace slots, and one Krylov solve: solve_pcg (the MG
// wrapped as a per-call fixed_spd preconditioner, which the reference's PCG
// requires) or solve_fgmres (the MG as given, or behind a volume-scaling
// wrapper with per-volume operator rows, the cold pressure pairing).  With a
// repeat count in the case file it also times the solve (wall clock of each
// repetition), for the CPU baseline.
//
// Reference interfaces used: include/hundun/v04_linear.hpp (MG plan, solve_pcg,
// workspaces, reduction engine), include/hundun/v04_parallel.hpp (HaloEngine),
// include/hundun/v04_mesh.hpp (geometry), src/solver_mg_detail.hpp (test
// hooks).
#include "hundun/v04_linear.hpp"
#include "hundun/v04_parallel.hpp"
#include "solver_mg_detail.hpp"

#include <mpi.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace hundun::v04;

namespace {

struct Reader {
  std::ifstream in;
  explicit Reader(const char* path) : in(path, std::ios::binary) {}
  template <class T>
  T get() {
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof v);
    return v;
  }
  template <class T>
  std::vector<T> vec(std::size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
    return v;
  }
};

struct Writer {
  std::ofstream out;
  explicit Writer(const char* path) : out(path, std::ios::binary) {}
  template <class T>
  void put(T v) {
    out.write(reinterpret_cast<const char*>(&v), sizeof v);
  }
  template <class T>
  void vec(const std::vector<T>& v) {
    out.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
  }
};

// Padded cell field with one ghost layer (FieldView over owned storage).
struct Field {
  std::vector<double> storage;
  FieldView view{};
  Field(Int3 cells, std::uint8_t components, FieldId id, StorageIdentity sid,
        std::uint8_t ghosts = 1) {
    const std::size_t nx = cells.x + 2 * ghosts, ny = cells.y + 2 * ghosts,
                      nz = cells.z + 2 * ghosts;
    storage.assign(nx * ny * nz * components, 0.0);
    view.base = storage.data() + ghosts + ghosts * nx + ghosts * nx * ny;
    view.interior = cells;
    view.ghosts = {ghosts, ghosts, ghosts};
    view.components = components;
    view.stride_y = nx;
    view.stride_z = nx * ny;
    view.component_stride = nx * ny * nz;
    view.field = id;
    view.revision = 1U;
    view.storage_identity = sid;
    view.revision_domain = 901U;
  }
};

ConstFaceFieldView face_view(const std::vector<double>& data, Int3 extents,
                             CartesianAxis axis, StorageIdentity sid) {
  ConstFaceFieldView v;
  v.base = data.data();
  v.extents = extents;
  v.stride_y = extents.x;
  v.stride_z = static_cast<std::size_t>(extents.x) * extents.y;
  v.axis = axis;
  v.storage_identity = sid;
  v.revision_domain = 902U;
  return v;
}

std::size_t cidx(Int3 n, int i, int j, int k) {
  return i + static_cast<std::size_t>(n.x) * (j + static_cast<std::size_t>(n.y) * k);
}

// Exact operator for the Krylov solve: the reference pressure operator
// (regular 7-point action with boundary-kind neighbours, then IBM
// decoration: solid rows identity, link faces removed), written here because
// its PISO host objects are outside the hot-path closure.
class HarnessOperator final : public LinearOperator {
 public:
  Int3 n{};
  int bc[6]{};
  const std::vector<double>* D{};
  const std::vector<double>* fx{};
  const std::vector<double>* fy{};
  const std::vector<double>* fz{};
  const std::vector<std::uint8_t>* act{};
  LinearIdentity identity{};

  LinearOperatorCertificate certificate() const noexcept override {
    return {identity, 77U, n, LinearOperatorClass::spd};
  }

  double nb(ConstFieldView x, int i, int j, int k, int axis, int dir) const {
    int c[3] = {i, j, k};
    const int e[3] = {n.x, n.y, n.z};
    c[axis] += dir;
    if (c[axis] >= 0 && c[axis] < e[axis]) return x.unchecked({c[0], c[1], c[2]}, 0);
    const int kind = bc[2 * axis + (dir > 0 ? 1 : 0)];
    if (kind == 0) return 0.0;
    if (kind == 1) return x.unchecked({i, j, k}, 0);
    c[axis] = dir < 0 ? e[axis] - 1 : 0;
    return x.unchecked({c[0], c[1], c[2]}, 0);
  }

  Status apply(FieldView xin, FieldView y) const noexcept override {
    const ConstFieldView x = as_const(xin);
    const Int3 xe{n.x + 1, n.y, n.z}, ye{n.x, n.y + 1, n.z}, ze{n.x, n.y, n.z + 1};
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i) {
          double v = (*D)[cidx(n, i, j, k)] * x.unchecked({i, j, k}, 0);
          v -= (*fx)[cidx(xe, i, j, k)] * nb(x, i, j, k, 0, -1);
          v -= (*fx)[cidx(xe, i + 1, j, k)] * nb(x, i, j, k, 0, 1);
          v -= (*fy)[cidx(ye, i, j, k)] * nb(x, i, j, k, 1, -1);
          v -= (*fy)[cidx(ye, i, j + 1, k)] * nb(x, i, j, k, 1, 1);
          v -= (*fz)[cidx(ze, i, j, k)] * nb(x, i, j, k, 2, -1);
          v -= (*fz)[cidx(ze, i, j, k + 1)] * nb(x, i, j, k, 2, 1);
          y.unchecked({i, j, k}, 0) = v;
        }
    if (act == nullptr || act->empty()) return {};
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i)
          if (!(*act)[cidx(n, i, j, k)])
            y.unchecked({i, j, k}, 0) = x.unchecked({i, j, k}, 0);
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i) {
          if (!(*act)[cidx(n, i, j, k)]) continue;
          for (int axis = 0; axis < 3; ++axis)
            for (int dir = -1; dir <= 1; dir += 2) {
              int c[3] = {i, j, k};
              const int e[3] = {n.x, n.y, n.z};
              c[axis] += dir;
              if (c[axis] < 0 || c[axis] >= e[axis]) {
                if (bc[2 * axis + (dir > 0)] != 2) continue;
                c[axis] = dir < 0 ? e[axis] - 1 : 0;
              }
              if ((*act)[cidx(n, c[0], c[1], c[2])]) continue;
              double face;
              if (axis == 0)
                face = (*fx)[cidx(xe, dir < 0 ? i : i + 1, j, k)];
              else if (axis == 1)
                face = (*fy)[cidx(ye, i, dir < 0 ? j : j + 1, k)];
              else
                face = (*fz)[cidx(ze, i, j, dir < 0 ? k : k + 1)];
              y.unchecked({i, j, k}, 0) +=
                  face * (x.unchecked({c[0], c[1], c[2]}, 0) -
                          x.unchecked({i, j, k}, 0));
            }
        }
    return {};
  }
};

// solve_pcg admits only a per-call fixed_spd preconditioner; this adapter
// presents the MG plan that way (the harness's choice, see the header note).
class PcgAdapter final : public LinearPreconditioner {
 public:
  NativeCartesianMgPlan* plan{};
  LinearIdentity identity{};
  LinearPreconditionerCertificate certificate() const noexcept override {
    return {identity, 78U, LinearPreconditionerClass::fixed_spd,
            LinearPreconditionerStatusScope::rank_local,
            LinearPreconditionerApplyLifecycle::per_call_checked};
  }
  Status apply(ConstFieldView in, FieldView out,
               std::uint32_t iteration) noexcept override {
    return plan->apply(in, out, iteration);
  }
};

// Operator rows per unit volume: the harness operator divided by the cell
// volume (w_x w_y) w_z (the cold pressure rows, src/solver_cold.hpp:417-470,
// are volume-normalised).
class VolumeRowsOperator final : public LinearOperator {
 public:
  const HarnessOperator* base{};
  const std::vector<double>* volume{};
  LinearOperatorCertificate certificate() const noexcept override {
    LinearOperatorCertificate c = base->certificate();
    c.operator_class = LinearOperatorClass::nonsymmetric;
    return c;
  }
  Status apply(FieldView x, FieldView y) const noexcept override {
    const Status st = base->apply(x, y);
    if (!st) return st;
    const Int3 n = base->n;
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i)
          y.unchecked({i, j, k}, 0) =
              y.unchecked({i, j, k}, 0) / (*volume)[cidx(n, i, j, k)];
    return {};
  }
};

// The MG applied to V r: the arithmetic of VolumeScaledPreconditioner
// (src/solver_cold.hpp:378-415), whose CartesianKernelPlan lies outside the
// compiled closure.
class VolumeScaledAdapter final : public LinearPreconditioner {
 public:
  NativeCartesianMgPlan* plan{};
  const std::vector<double>* volume{};
  FieldView scratch{};
  Int3 n{};
  LinearPreconditionerCertificate certificate() const noexcept override {
    return plan->certificate();
  }
  Status prepare_batch(const LinearPreconditionerBatchDescriptor& descriptor,
                       LinearPreconditionerBatchTicket& ticket) noexcept override {
    return plan->prepare_batch(descriptor, ticket);
  }
  Status apply_prepared(ConstFieldView residual, FieldView correction,
                        std::uint32_t iteration,
                        const LinearPreconditionerBatchTicket& ticket) noexcept override {
    scale(residual);
    return plan->apply_prepared(as_const(scratch), correction, iteration,
                                ticket);
  }
  Status apply(ConstFieldView residual, FieldView correction,
               std::uint32_t iteration) noexcept override {
    scale(residual);
    return plan->apply(as_const(scratch), correction, iteration);
  }

 private:
  void scale(ConstFieldView residual) {
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i)
          scratch.unchecked({i, j, k}, 0) =
              residual.unchecked({i, j, k}, 0) * (*volume)[cidx(n, i, j, k)];
  }
};

std::uint8_t mask_of(CoarseningKind k) {
  switch (k) {
    case CoarseningKind::full_xyz: return 7;
    case CoarseningKind::semi_xy: return 3;
    case CoarseningKind::semi_xz: return 5;
    case CoarseningKind::semi_yz: return 6;
    case CoarseningKind::x_only: return 1;
    case CoarseningKind::y_only: return 2;
    case CoarseningKind::z_only: return 4;
  }
  return 0;
}

int fail(const char* what, Status s) {
  std::cerr << "ref_driver: " << what << " failed code="
            << static_cast<int>(s.code) << " detail=" << s.detail << "\n";
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: hf_ref_driver <case.bin> <out.bin>\n";
    return 2;
  }
  MPI_Init(&argc, &argv);
  int rc = 0;
  // HF_REF_TIMING_ONLY: the CPU-baseline mode -- compile and time the Krylov
  // solve only; the level diagonals, the single application with its
  // workspace slots and the solution are not written.
  const bool timing_only = std::getenv("HF_REF_TIMING_ONLY") != nullptr;
  {
    Reader in(argv[1]);
    char magic[8];
    in.in.read(magic, 8);
    if (std::memcmp(magic, "HFREF002", 8) != 0) {
      std::cerr << "bad case file\n";
      MPI_Finalize();
      return 2;
    }
    const Int3 n{in.get<std::int32_t>(), in.get<std::int32_t>(),
                 in.get<std::int32_t>()};
    int bc[6];
    for (int& v : bc) v = in.get<std::int32_t>();
    MgHierarchyPolicy policy;
    policy.anisotropy_threshold = in.get<double>();
    policy.pre_sweeps = static_cast<std::uint8_t>(in.get<std::int32_t>());
    policy.post_sweeps = static_cast<std::uint8_t>(in.get<std::int32_t>());
    policy.maximum_levels = static_cast<std::uint8_t>(in.get<std::int32_t>());
    policy.coarse_sweeps = static_cast<std::uint16_t>(in.get<std::int32_t>());
    policy.minimum_coarse_extent =
        static_cast<std::uint8_t>(in.get<std::int32_t>());
    policy.cycle = in.get<std::int32_t>() ? MgCycleKind::f_cycle
                                          : MgCycleKind::v_cycle;
    const bool null_space = in.get<std::int32_t>() != 0;
    const bool unit_linear = in.get<std::int32_t>() != 0;
    const bool has_activity = in.get<std::int32_t>() != 0;
    const double pcg_rtol = in.get<double>();
    const std::int32_t pcg_maxit = in.get<std::int32_t>();
    const std::int32_t pcg_interval = in.get<std::int32_t>();
    policy.point_smoother = in.get<std::int32_t>()
                                ? MgPointSmootherKind::chebyshev_jacobi
                                : MgPointSmootherKind::red_black;
    policy.chebyshev_lower_spectrum_fraction = in.get<double>();
    policy.diagonal_shortcut_maximum_ratio = in.get<double>();
    const std::int32_t solver = in.get<std::int32_t>();  // 0 pcg, 1 fgmres
    const std::int32_t restart = in.get<std::int32_t>();
    const bool volume_rows = in.get<std::int32_t>() != 0;
    const std::int32_t repeats = in.get<std::int32_t>();
    const std::size_t cells = static_cast<std::size_t>(n.x) * n.y * n.z;
    const Int3 xe{n.x + 1, n.y, n.z}, ye{n.x, n.y + 1, n.z}, ze{n.x, n.y, n.z + 1};
    const std::size_t nfx = static_cast<std::size_t>(xe.x) * xe.y * xe.z;
    const std::size_t nfy = static_cast<std::size_t>(ye.x) * ye.y * ye.z;
    const std::size_t nfz = static_cast<std::size_t>(ze.x) * ze.y * ze.z;
    std::vector<double> faces[3] = {in.vec<double>(n.x + 1),
                                    in.vec<double>(n.y + 1),
                                    in.vec<double>(n.z + 1)};
    std::vector<double> D = in.vec<double>(cells);
    std::vector<double> bx = in.vec<double>(nfx);
    std::vector<double> by = in.vec<double>(nfy);
    std::vector<double> bz = in.vec<double>(nfz);
    std::vector<std::uint8_t> ca, fxa, fya, fza;
    if (has_activity) {
      ca = in.vec<std::uint8_t>(cells);
      fxa = in.vec<std::uint8_t>(nfx);
      fya = in.vec<std::uint8_t>(nfy);
      fza = in.vec<std::uint8_t>(nfz);
    }
    std::vector<double> r = in.vec<double>(cells);
    std::vector<double> b = in.vec<double>(cells);
    if (!in.in) {
      std::cerr << "truncated case file\n";
      MPI_Finalize();
      return 2;
    }

    CartesianMeshSpec mesh;
    mesh.kind = GeometryKind::runtime_axes_v1;
    // runtime axes carry REAL4 face coordinates (src/mesh_cartesian.cpp:
    // 324-354); the case files already hold float-representable faces
    mesh.axes_file = "runtime_axes";
    mesh.has_exact_cells = true;
    mesh.exact_cells = n;
    for (int a = 0; a < 3; ++a) mesh.runtime_faces[a] = faces[a];
    mesh.lower = {faces[0].front(), faces[1].front(), faces[2].front()};
    mesh.upper = {faces[0].back(), faces[1].back(), faces[2].back()};
    mesh.limits = {1000000000U, 1ULL << 40U};
    CartesianGeometryPlan geometry;
    MeshPatch patch;
    Status st = CartesianGeometryCompiler::compile(MPI_COMM_SELF, mesh,
                                                   GeometryBudget{}, geometry,
                                                   patch);
    if (!st) return fail("geometry", st);

    const MgBoundaryKind kinds[3] = {MgBoundaryKind::dirichlet,
                                     MgBoundaryKind::neumann,
                                     MgBoundaryKind::periodic};
    NativeCartesianMgSpec spec;
    spec.communicator = MPI_COMM_SELF;
    spec.geometry = &geometry;
    spec.patch = patch;
    spec.boundaries = {kinds[bc[0]], kinds[bc[1]], kinds[bc[2]],
                       kinds[bc[3]], kinds[bc[4]], kinds[bc[5]]};
    spec.null_space = null_space ? MgNullSpace::constant : MgNullSpace::none;
    spec.policy = policy;
    spec.correction_scaling = unit_linear
                                  ? MgCorrectionScaling::unit_linear
                                  : MgCorrectionScaling::residual_minimizing;
    spec.identity = {301U, 302U, 303U, 304U, 305U};
    spec.coefficients = {1U, 401U, 0.0};
    if (has_activity) {
      spec.activity.cells = {ca.data(), ca.size()};
      spec.activity.x_faces = {fxa.data(), fxa.size()};
      spec.activity.y_faces = {fya.data(), fya.size()};
      spec.activity.z_faces = {fza.data(), fza.size()};
      spec.activity.local_fingerprint = 611U;
      spec.activity.collective_fingerprint = 612U;
    }
    MgWorkspaceRequirements wreq;
    st = make_mg_workspace_requirements(MPI_COMM_SELF, geometry, patch, policy,
                                        81U, wreq);
    if (!st) return fail("mg workspace requirements", st);
    Field vectors(wreq.arena_shape, 1U, 10U, 500U, 0U);
    MgWorkspace workspace;
    ReductionEngine reductions;
    st = MgWorkspace::bind(wreq, vectors.view, workspace);
    if (!st) return fail("mg workspace bind", st);
    st = ReductionEngine::compile(MPI_COMM_SELF, ReductionMode::mpi_allreduce,
                                  64U, reductions);
    if (!st) return fail("reductions", st);
    const HaloFieldSpec halo_fields[1] = {{10U, 1U, 1U}};
    const HaloTopology topology{bc[0] == 2, bc[2] == 2, bc[4] == 2};
    HaloEngine halo;
    st = halo.reserve(MPI_COMM_SELF, wreq.levels[0].patch, {halo_fields, 1U},
                      topology);
    if (!st) return fail("halo", st);
    std::vector<HaloEngine> coarse_halos(wreq.level_count - 1U);
    std::vector<HaloEngine*> coarse_ptrs(coarse_halos.size());
    for (std::size_t l = 1; l < wreq.level_count; ++l) {
      st = coarse_halos[l - 1].reserve(MPI_COMM_SELF, wreq.levels[l].patch,
                                       {halo_fields, 1U}, topology);
      if (!st) return fail("coarse halo", st);
      coarse_ptrs[l - 1] = &coarse_halos[l - 1];
    }
    const MgRuntimeServices services{&halo, &reductions, &workspace,
                                     {coarse_ptrs.data(), coarse_ptrs.size()}};
    Field diag(n, 1U, 1U, 700U, 0U);
    for (std::size_t c = 0; c < cells; ++c) diag.view.base[c] = D[c];
    const MgCoefficientViews coeffs{
        as_const(diag.view), face_view(bx, xe, CartesianAxis::x, 710U),
        face_view(by, ye, CartesianAxis::y, 711U),
        face_view(bz, ze, CartesianAxis::z, 712U)};
    // the Krylov workspace comes first: its fingerprint is part of the
    // identity that the solver, the operator and the preconditioner share
    LinearWorkspaceRequirements lreq;
    const bool fgmres = solver == 1;
    st = make_linear_workspace_requirements(
        fgmres ? LinearAlgorithm::fgmres : LinearAlgorithm::pcg, n, 1U,
        fgmres ? static_cast<std::uint32_t>(restart) : 0U,
        ReductionMode::mpi_allreduce, 91U, lreq);
    if (!st) return fail("linear workspace requirements", st);
    Field lvec(n, lreq.vector_slots, 40U, 740U, 1U);
    Field lscalar({static_cast<std::int32_t>(lreq.scalar_doubles), 1, 1}, 1U,
                  41U, 740U, 0U);
    SolverWorkspace lws;
    st = SolverWorkspace::bind(lreq, lvec.view, lscalar.view, lws);
    if (!st) return fail("solver workspace bind", st);
    LinearIdentity identity{301U, 302U, 303U, lws.fingerprint(), 305U};
    spec.operator_class =
        MgOperatorClass::symmetric_diagonally_dominant_m_matrix;
    spec.identity = identity;
    NativeCartesianMgPlan plan;
    const double compile0 = MPI_Wtime();
    st = NativeCartesianMgPlan::compile(spec, services, coeffs, plan);
    const double compile_seconds = MPI_Wtime() - compile0;
    if (!st) return fail("mg compile", st);

    Writer out(argv[2]);
    out.out.write("HFOUT002", 8);
    const std::size_t levels = plan.level_count();
    out.put<std::int32_t>(static_cast<std::int32_t>(levels));
    out.put<std::int32_t>(plan.uses_diagonal_preconditioner() ? 1 : 0);
    std::vector<std::size_t> level_cells(levels);
    for (std::size_t l = 0; l < levels; ++l) {
      MgLevelView v;
      plan.level(l, v);
      out.put<std::int32_t>(v.local_shape.x);
      out.put<std::int32_t>(v.local_shape.y);
      out.put<std::int32_t>(v.local_shape.z);
      out.put<std::int32_t>(wreq.levels[l].coarsen_axis_mask);
      out.put<std::int32_t>(v.line_axis_mask);
      level_cells[l] = static_cast<std::size_t>(v.local_shape.x) *
                       v.local_shape.y * v.local_shape.z;
    }
    for (std::size_t l = 0; l < levels && !timing_only; ++l) {
      std::vector<double> d(level_cells[l]);
      for (std::size_t c = 0; c < level_cells[l]; ++c)
        d[c] = detail::mg_level_diagonal_for_test(plan, l, c);
      out.vec(d);
    }
    // one preconditioner application
    if (!timing_only) {
    Field rin(n, 1U, 30U, 730U, 0U), zout(n, 1U, 31U, 731U, 0U);
    for (std::size_t c = 0; c < cells; ++c) rin.view.base[c] = r[c];
    st = plan.apply(as_const(rin.view), zout.view, 0U);
    if (!st) return fail("mg apply", st);
    std::vector<double> z(zout.view.base, zout.view.base + cells);
    out.vec(z);
    }
    // every level's solution / rhs / residual slot after the application
    for (std::size_t l = 0; l < levels && !timing_only; ++l) {
      for (MgWorkspaceSlot slot : {MgWorkspaceSlot::solution,
                                   MgWorkspaceSlot::rhs,
                                   MgWorkspaceSlot::residual}) {
        const FieldView v = workspace.level(l, slot);
        std::vector<double> values;
        values.reserve(level_cells[l]);
        for (int k = 0; k < v.interior.z; ++k)
          for (int j = 0; j < v.interior.y; ++j)
            for (int i = 0; i < v.interior.x; ++i)
              values.push_back(v.unchecked({i, j, k}, 0));
        out.vec(values);
      }
    }
    HarnessOperator op;
    op.n = n;
    std::memcpy(op.bc, bc, sizeof bc);
    op.D = &D;
    op.fx = &bx;
    op.fy = &by;
    op.fz = &bz;
    op.act = &ca;
    op.identity = identity;
    PcgAdapter pc;
    pc.plan = &plan;
    pc.identity = identity;
    std::vector<double> volume(cells);
    for (int k = 0; k < n.z; ++k)
      for (int j = 0; j < n.y; ++j)
        for (int i = 0; i < n.x; ++i)
          volume[cidx(n, i, j, k)] =
              (faces[0][i + 1] - faces[0][i]) * (faces[1][j + 1] - faces[1][j]) *
              (faces[2][k + 1] - faces[2][k]);
    VolumeRowsOperator vop;
    vop.base = &op;
    vop.volume = &volume;
    Field vscratch(n, 1U, 52U, 752U, 1U);
    VolumeScaledAdapter vpc;
    vpc.plan = &plan;
    vpc.volume = &volume;
    vpc.scratch = vscratch.view;
    vpc.n = n;
    Field rhs(n, 1U, 50U, 750U, 0U), sol(n, 1U, 51U, 751U, 0U);
    for (std::size_t c = 0; c < cells; ++c) rhs.view.base[c] = b[c];
    LinearSolveInvocation inv;
    inv.rhs = as_const(rhs.view);
    inv.solution = sol.view;
    inv.expected_identity = identity;
    inv.control.absolute_tolerance = 0.0;
    inv.control.relative_tolerance = pcg_rtol;
    inv.control.maximum_iterations = static_cast<std::uint32_t>(pcg_maxit);
    inv.control.true_residual_interval = static_cast<std::uint32_t>(pcg_interval);
    if (fgmres) inv.control.restart = static_cast<std::uint32_t>(restart);
    auto solve = [&]() {
      for (std::size_t c = 0; c < cells; ++c) sol.view.base[c] = 0.0;
      if (!fgmres) return solve_pcg(op, pc, inv, lws, reductions);
      if (volume_rows) return solve_fgmres(vop, vpc, inv, lws, reductions);
      return solve_fgmres(op, plan, inv, lws, reductions);
    };
    // the timing-only mode times every solve it runs (no untimed warm-up:
    // the first timed solve gives the result)
    std::vector<double> seconds;
    const double first0 = MPI_Wtime();
    const LinearSolveResult res = solve();
    if (timing_only) seconds.push_back(MPI_Wtime() - first0);
    for (std::int32_t rep = timing_only ? 1 : 0; rep < repeats; ++rep) {
      const double t0 = MPI_Wtime();
      const LinearSolveResult again = solve();
      seconds.push_back(MPI_Wtime() - t0);
      if (again.iterations != res.iterations) rc = 4;
    }
    out.put<std::int32_t>(static_cast<std::int32_t>(res.status.code));
    out.put<std::int32_t>(static_cast<std::int32_t>(res.termination));
    out.put<std::int32_t>(static_cast<std::int32_t>(res.iterations));
    out.put<double>(res.initial_true_residual);
    out.put<double>(res.final_true_residual);
    out.put<double>(res.rhs_norm);
    out.put<std::int64_t>(static_cast<std::int64_t>(res.operator_applies));
    out.put<std::int64_t>(static_cast<std::int64_t>(res.preconditioner_applies));
    out.put<std::int64_t>(static_cast<std::int64_t>(res.norm_breakdown_restarts));
    if (!timing_only) {
      std::vector<double> x(sol.view.base, sol.view.base + cells);
      out.vec(x);
    }
    out.put<std::int32_t>(static_cast<std::int32_t>(seconds.size()));
    out.vec(seconds);
    if (timing_only) out.put<double>(compile_seconds);
    if (!out.out) rc = 3;
    std::cerr << "ref_driver: levels=" << levels << " solver="
              << (fgmres ? "fgmres" : "pcg") << " status="
              << static_cast<int>(res.status.code) << "/" << res.status.detail << " iterations="
              << res.iterations << " final=" << res.final_true_residual
              << "\n";
  }
  MPI_Finalize();
  return rc;
}
