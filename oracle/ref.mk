# Builds lpm-v2's own Fortran modules of the direct-sum path, from the
# sources where they lie under /root/reference/src (nothing is copied), with
# amdflang and the image's MPICH (/opt/conda: mpif.h, libmpifort, libmpi),
# and links the golden-vector driver tests/golden/ref_golden.f90 against
# them.  Outputs go to oracle/_ref/ only.  Test infrastructure: nothing in
# the library or the benchmark's timed path uses it.
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

all: $(OUT)/ref_golden

$(OUT)/mod:
	mkdir -p $(OUT)/mod

# one module at a time, in order (each needs the .mod files of the ones before)
$(OUT)/lpm_ref_modules.stamp: | $(OUT)/mod
	set -e; for m in $(MODS); do $(AMDFLANG) $(FFLAGS) -c $(REF)/$$m.f90 -o $(OUT)/$$m.o; done
	touch $@

$(OUT)/ref_golden: tests/golden/ref_golden.f90 $(OUT)/lpm_ref_modules.stamp
	$(AMDFLANG) $(FFLAGS) -c tests/golden/ref_golden.f90 -o $(OUT)/ref_golden.o
	$(AMDFLANG) -o $@ $(OUT)/ref_golden.o $(OBJS) -L$(MPI)/lib -lmpifort -lmpi -Wl,-rpath,$(MPI)/lib

clean:
	rm -rf $(OUT)

.PHONY: all clean
