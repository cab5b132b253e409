# Builds the reference's multigrid pressure solver, its velocity projection
# (CORE_FRAC_STEP_ROUTINES.F90), the advection-diffusion coefficients
# (CORE_SET_SOLVERS.F90) and the immersed-boundary geometry routines
# (BOUNDARY_SET_ARCLENGTH_NORM.F90, BOUNDARY_IBLANK_FAST.F90,
# GCM_SET_BOUNDARY.F90, GCM_MEMORY_ALLOCATE.F90, GCM_BODYINTERCEPT_VALUES.F90)
# with the marker stresses (UTIL_DRAG_LIFT.F90) from their own sources, where
# they lie in the reference tree, unmodified, into oracle/_ref/ (git-ignored),
# plus ref_harness (oracle/ref_harness.f90) that drives it.  Test
# infrastructure only: the tests compare it with the C oracle.
#
#   make -f oracle/ref.mk REF=/root/reference
#
# Compiler: flang from ROCm; MPI: the MPICH of the conda tree (mpif.h and the
# Fortran bindings); LAPACK (DGETRF/DGETRI of GCM_Calc_VanMatrixDN): the MKL
# of the conda tree, as the reference's makefile links MKL (-qmkl).  The reference's own makefile passes no -D.
# -DPE12=1PE12 is this recipe's: the preprocessor rewrites the Intel-only
# edit descriptor PE12.5 of the FORMAT statement at CORE_MG.F90:1041 (an
# infoconv print) into the standard 1PE12.5, which flang accepts.  It changes
# a format only; no arithmetic is touched.
#
# GCM_GHOSTCELL_UPDATE.F90 does not compile with flang (Intel-only 'P' edit
# descriptors at :349 and :734), and BOUNDARY_MARKER_INIT.F90 is not needed, so
# the objects leave four symbols undefined: gcm_ghostcell_pressure_,
# gcm_enforce_p_compatibility_ and correct_ghost_massflux_ (CORE_POISSON.F90:
# solve_poisson and rhs_poisson) and read_inputs_body_ (CORE_PARALLEL.F90:
# MPI_Initialize).  None of them is on a path the harness calls.  The objects
# go into a shared library, whose undefined functions the dynamic linker binds
# lazily at their first call, so a call into one of them would stop the
# program with an "undefined symbol" error rather than run anything else.
REF    ?= /root/reference
HERE   := $(dir $(lastword $(MAKEFILE_LIST)))
OUT    := $(HERE)_ref
FLANG  ?= /opt/rocm/llvm/bin/flang
MPI    ?= /opt/conda
FFLAGS := -O2 -fPIC -cpp -DPE12=1PE12 -I$(MPI)/include
SRCS   := AMODULES CORE_PARALLEL CORE_TDMA CORE_SET_OUTER_BC CORE_GRID CORE_MEMORY_ALLOCATE CORE_MG CORE_POISSON CORE_FRAC_STEP_ROUTINES CORE_SET_SOLVERS \
          BOUNDARY_SET_ARCLENGTH_NORM BOUNDARY_IBLANK_FAST GCM_SET_BOUNDARY GCM_MEMORY_ALLOCATE \
          GCM_BODYINTERCEPT_VALUES UTIL_DRAG_LIFT
OBJS   := $(addprefix $(OUT)/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/ref_harness

$(OUT)/AMODULES.o: $(REF)/AMODULES.F90
	mkdir -p $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o AMODULES.o

$(OUT)/%.o: $(REF)/%.F90 $(OUT)/AMODULES.o
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o $*.o

$(OUT)/libref_mg.so: $(OBJS)
	$(FLANG) -shared -o $@ $(OBJS) -L$(MPI)/lib -lmpifort -lmpi -lmkl_rt -Wl,-rpath,$(MPI)/lib -Wl,-z,lazy

$(OUT)/ref_harness: $(HERE)ref_harness.f90 $(OUT)/libref_mg.so
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $(abspath $(HERE)ref_harness.f90) -o ref_harness.o
	$(FLANG) -o $@ $(OUT)/ref_harness.o -L$(OUT) -lref_mg -L$(MPI)/lib -lmpifort -lmpi -lmkl_rt \
	    -Wl,-rpath,$(abspath $(OUT)) -Wl,-rpath,$(MPI)/lib -Wl,--allow-shlib-undefined -Wl,-z,lazy

clean:
	rm -rf $(OUT)

.PHONY: all clean
