# Builds the reference's own hot-path closure (13 translation units of the
# reference tree, compiled where they lie) plus ref_driver.cpp into _ref/.
# Test infrastructure only: _ref/ is git-ignored and never linked into the
# product.  Needs the reference tree and MPICH (headers/libmpi from
# /opt/conda); the reference CMake is not used.
REF      ?= /root/reference/versions/v0.4
MPI_ROOT ?= /opt/conda
CXX      ?= g++
OUT      := _ref
OBJDIR   := $(OUT)/obj
CXXFLAGS := -std=c++17 -O2 -ffp-contract=off -DHUNDUN_V04_ENABLE_TEST_ACCESS=1 \
            -I$(REF)/include -I$(REF)/src -I$(MPI_ROOT)/include -w
UNITS    := solver_mg solver_mg_workspace solver_krylov solver_reduction \
            parallel_halo solver_linear_state mesh_cartesian mesh_focus \
            mesh_decomposition core_field_registry core_status core_revision \
            core_arena
OBJS     := $(addprefix $(OBJDIR)/,$(addsuffix .o,$(UNITS))) $(OBJDIR)/ref_driver.o
MPILIBS  := libmpi.so.12 libgfortran.so.4 libquadmath.so.0 libgomp.so.1

all: $(OUT)/hf_ref_driver

$(OBJDIR)/%.o: $(REF)/src/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJDIR)/ref_driver.o: ref_driver.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJDIR):
	mkdir -p $(OBJDIR) $(OUT)/mpilib

# libmpi from the conda tree is linked through a private directory so the
# conda libstdc++ never shadows the system one at run time
$(OUT)/mpilib/.linked: | $(OBJDIR)
	for l in $(MPILIBS); do ln -sf $(MPI_ROOT)/lib/$$l $(OUT)/mpilib/$$l; done
	ln -sf $(MPI_ROOT)/lib/libmpi.so.12 $(OUT)/mpilib/libmpi.so
	touch $@

$(OUT)/hf_ref_driver: $(OBJS) $(OUT)/mpilib/.linked
	$(CXX) -o $@ $(OBJS) -L$(OUT)/mpilib -lmpi -Wl,-rpath,'$$ORIGIN/mpilib' \
	    -Wl,-rpath-link,$(OUT)/mpilib

clean:
	rm -rf $(OUT)

.PHONY: all clean
