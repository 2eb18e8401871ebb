# Builds lpm-v2's own Fortran modules of the direct-sum path, from the
# sources where they lie under /root/reference/src (nothing is copied), with
# amdflang and the image's MPICH (/opt/conda: mpif.h, libmpifort, libmpi),
# and links the golden-vector driver tests/golden/ref_golden.f90 and the CPU
# baseline driver oracle/ref_cpu_bench.f90 against them.  Outputs go to
# oracle/_ref/ only.  Test infrastructure: nothing in the library or the
# benchmark's timed path uses it.
#
#   make -f oracle/ref.mk          (from the repository root)
#
# -O2 is the reference's own choice for compilers other than GNU and Intel
# (CMakeLists.txt:36-40).
REF      ?= /root/reference/src
OUT      ?= oracle/_ref
AMDFLANG ?= /opt/rocm/bin/amdflang
MPI      ?= /opt/conda
FFLAGS   ?= -cpp -O2 -I$(MPI)/include -J $(OUT)/mod
# the modules of the path, in dependency order
MODS     := TypeDefs Utilities STDIntVector OutputWriter Logger SphereGeometry PlaneGeometry Particles Edges \
            Faces Field PolyMesh2d MPISetup SphereBVE SphereBVESolver PlanarIncompressible \
            PlaneIncompressibleSolver PSEDirectSum
OBJS     := $(addprefix $(OUT)/,$(addsuffix .o,$(MODS)))

all: $(OUT)/ref_golden $(OUT)/ref_gpu_timestep $(OUT)/ref_cpu_bench

# The GPU-backed SphereBVESolverModule (lpm-v2_amd/fortran/SphereBVESolver_lpm.f90)
# linked with the reference's other modules in place of its SphereBVESolver,
# driven by the reference's own call sequence (tests/golden/ref_gpu_timestep.f90).
# Its .mod files go to a directory of their own, searched first, since the
# module name is the reference's.
LPMLIB   ?= lpm-v2_amd
GPUMOD   := $(OUT)/mod_gpu
GPUOBJS  := $(filter-out $(OUT)/SphereBVESolver.o,$(OBJS))

$(OUT)/ref_gpu_timestep: tests/golden/ref_gpu_timestep.f90 lpm-v2_amd/fortran/SphereBVESolver_lpm.f90 \
                         lpm-v2_amd/fortran/lpm_direct_f.f90 $(OUT)/lpm_ref_modules.stamp
	mkdir -p $(GPUMOD)
	$(AMDFLANG) -cpp -O2 -J $(GPUMOD) -c lpm-v2_amd/fortran/lpm_direct_f.f90 -o $(GPUMOD)/lpm_direct_f.o
	$(AMDFLANG) -cpp -O2 -I$(MPI)/include -J $(GPUMOD) -I $(OUT)/mod -c lpm-v2_amd/fortran/SphereBVESolver_lpm.f90 \
		-o $(GPUMOD)/SphereBVESolver_lpm.o
	$(AMDFLANG) -cpp -O2 -I$(MPI)/include -I $(GPUMOD) -I $(OUT)/mod -J $(GPUMOD) -c tests/golden/ref_gpu_timestep.f90 \
		-o $(GPUMOD)/ref_gpu_timestep.o
	$(AMDFLANG) -o $@ $(GPUMOD)/ref_gpu_timestep.o $(GPUMOD)/SphereBVESolver_lpm.o $(GPUMOD)/lpm_direct_f.o \
		$(GPUOBJS) -L$(MPI)/lib -lmpifort -lmpi -L$(LPMLIB) -llpm_direct -Wl,-rpath,$(MPI)/lib \
		-Wl,-rpath,'$$ORIGIN/../../$(LPMLIB)'

$(OUT)/mod:
	mkdir -p $(OUT)/mod

# one module at a time, in order (each needs the .mod files of the ones before)
$(OUT)/lpm_ref_modules.stamp: | $(OUT)/mod
	set -e; for m in $(MODS); do $(AMDFLANG) $(FFLAGS) -c $(REF)/$$m.f90 -o $(OUT)/$$m.o; done
	touch $@

$(OUT)/ref_golden: tests/golden/ref_golden.f90 $(OUT)/lpm_ref_modules.stamp
	$(AMDFLANG) $(FFLAGS) -c tests/golden/ref_golden.f90 -o $(OUT)/ref_golden.o
	$(AMDFLANG) -o $@ $(OUT)/ref_golden.o $(OBJS) -L$(MPI)/lib -lmpifort -lmpi -Wl,-rpath,$(MPI)/lib

# The reference's own Biot-Savart loop timed on the host cores (bench.py's
# CPU baseline, one MPI rank per core).
$(OUT)/ref_cpu_bench: oracle/ref_cpu_bench.f90 $(OUT)/lpm_ref_modules.stamp
	$(AMDFLANG) $(FFLAGS) -c oracle/ref_cpu_bench.f90 -o $(OUT)/ref_cpu_bench.o
	$(AMDFLANG) -o $@ $(OUT)/ref_cpu_bench.o $(OBJS) -L$(MPI)/lib -lmpifort -lmpi -Wl,-rpath,$(MPI)/lib

clean:
	rm -rf $(OUT)

.PHONY: all clean
