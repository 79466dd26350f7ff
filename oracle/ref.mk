# Recipe for oracle/_ref/: the reference's own steady-transport hot-path sources, compiled in
# place from the read-only reference tree, linked with the one-rank harness of this directory.
# Used only in the development container (the reference is not present anywhere else).
#   make -f oracle/ref.mk REF=/root/reference
# Strict IEEE flags (the reference Makefile uses -Ofast; fast-math would only perturb the last
# bits of the golden vectors).
REF     ?= /root/reference
MPI     ?= /opt/conda
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
CF      := $(REF)/00_CommonFiles
ST      := $(REF)/01_SS_Transport_Equation/2D/01_CommonFiles
CFLAGS  := -O2 -fno-fast-math -ffp-contract=off -w -DSSTranspEquation2D -I$(ST) -I$(MPI)/include

ST_SRC := $(ST)/Fill_LM.c $(ST)/Permutation_of_LM.c $(ST)/csr_Initialization.c \
          $(ST)/Build_K_F_SUPG.c $(ST)/F_assembly.c $(ST)/CAU_ShockCapture.c \
          $(ST)/h_shock_2sqrtArea.c $(ST)/ebe_assembly.c $(ST)/csr_assembly.c $(ST)/setzeros.c
CF_SRC := $(CF)/Allocation_Operations/mycalloc.c \
          $(CF)/BLAS_Operations/daxpy.c $(CF)/BLAS_Operations/dcopy.c $(CF)/BLAS_Operations/ddot.c \
          $(CF)/BLAS_Operations/dscal.c $(CF)/BLAS_Operations/dzero.c \
          $(CF)/MatrixVector_Operations/EBE/ebemv.c $(CF)/MatrixVector_Operations/CSR/csrmv.c \
          $(CF)/Reordering/reordering.c $(CF)/Reordering/symrcm.c \
          $(CF)/Solvers_and_Preconditioners/NO_precond.c \
          $(CF)/Solvers_and_Preconditioners/NO_precond_setup.c \
          $(CF)/Solvers_and_Preconditioners/CSR/ILUp_precond.c \
          $(CF)/Solvers_and_Preconditioners/CSR/ILUp_precond_setup.c \
          $(CF)/Solvers_and_Preconditioners/pgmres.c

NS      := $(REF)/04_SS_NavierStokes_Equations/2D/01_CommonFiles
NS_SRC  := $(NS)/Build_K_F_SUPG_PSPG.c $(NS)/eval_U.c $(NS)/setzeros.c $(NS)/ebe_assembly.c \
           $(NS)/csr_assembly.c $(NS)/Fill_LM.c $(NS)/csr_Initialization.c
NS_CF   := $(CF)/Allocation_Operations/mycalloc.c $(CF)/BLAS_Operations/dzero.c \
           $(CF)/BLAS_Operations/ddot.c $(CF)/Reordering/reordering.c $(CF)/Reordering/symrcm.c

EU      := $(REF)/03_Euler_Equations/2D/01_CommonFiles
EU_SRC  := $(EU)/Build_M_K_F_SUPG.c $(EU)/eval_U_dU.c $(EU)/Ax_Ay_calculations.c $(EU)/Delta_CAU.c \
           $(EU)/setzeros.c $(EU)/ebe_assembly.c $(EU)/csr_assembly.c $(EU)/Fill_LM.c $(EU)/csr_Initialization.c \
           $(CF)/Solvers_and_Preconditioners/EBE/BlockDiag_precond_EBE_setup.c \
           $(CF)/Solvers_and_Preconditioners/BlockDiag_precond.c $(CF)/Solvers_and_Preconditioners/NO_precond.c \
           $(CF)/Solvers_and_Preconditioners/pgmres.c $(CF)/MatrixVector_Operations/EBE/ebemvNDOF4.c \
           $(CF)/BLAS_Operations/daxpy.c $(CF)/BLAS_Operations/dcopy.c $(CF)/BLAS_Operations/dscal.c

all: $(OUT)/st_harness $(OUT)/st_harness_ofast $(OUT)/st_harness_femgpu $(OUT)/ns_harness $(OUT)/eu_harness \
     $(OUT)/meshGenerator

# The structured mesh generator is a self-contained C program.
$(OUT)/meshGenerator: $(REF)/Mesh_Generator/Structured_Mesh/meshGenerator.c
	mkdir -p $(OUT)
	gcc -O2 -w -o $@ $< -lm

# mc73_fiedler (HSL_MC73, Spectral ordering only) is referenced by reordering.c but never
# called with SYMRCM/NOT; its prebuilt objects are not used, so the symbol stays unresolved.
$(OUT)/st_harness: $(HERE)ref_harness_st.c $(ST_SRC) $(CF_SRC)
	mkdir -p $(OUT)
	gcc $(CFLAGS) -o $@ $(HERE)ref_harness_st.c $(ST_SRC) $(CF_SRC) \
	    -L$(MPI)/lib -Wl,-rpath,$(MPI)/lib -lmpi -lm \
	    -Wl,--unresolved-symbols=ignore-in-object-files

# The same harness built with the reference Makefile's optimisation flags (-Ofast -march=native,
# 01_SS_Transport_Equation/2D/01_CommonFiles/Makefile:4; its -Ddebug only adds progress prints):
# the timing baseline of the reference CPU path (oracle/cpu_ratio.py), not a fixture source.
$(OUT)/st_harness_ofast: $(HERE)ref_harness_st.c $(ST_SRC) $(CF_SRC)
	mkdir -p $(OUT)
	gcc -Ofast -march=native -w -DSSTranspEquation2D -I$(ST) -I$(MPI)/include -o $@ $(HERE)ref_harness_st.c \
	    $(ST_SRC) $(CF_SRC) -L$(MPI)/lib -Wl,-rpath,$(MPI)/lib -lmpi -lm \
	    -Wl,--unresolved-symbols=ignore-in-object-files

# The same harness with the reference-side binding of the library: the reference's assembly and
# structure code around parallel_fem_codes_amd/compat/femgpu_compat.c and libfemgpu.so, which
# replace ebemv, csrmv, ILUp_precond(_setup), NO_precond_setup and pgmres (tests/test_gpu_shim.py
# runs it on the GPU).
LIBDIR  := $(abspath $(HERE)../parallel_fem_codes_amd)
ROCM    ?= /opt/rocm
SHIM_CF := $(filter-out %/ebemv.c %/csrmv.c %/ILUp_precond.c %/ILUp_precond_setup.c %/pgmres.c \
             %/NO_precond_setup.c,$(CF_SRC))
$(OUT)/st_harness_femgpu: $(HERE)ref_harness_st.c $(ST_SRC) $(SHIM_CF) $(HERE)../parallel_fem_codes_amd/compat/femgpu_compat.c \
                         $(LIBDIR)/libfemgpu.so
	mkdir -p $(OUT)
	gcc $(CFLAGS) -DFEMGPU_SHIM -fgnu89-inline -I$(CF) -I$(HERE)../include -I$(ROCM)/include -D__HIP_PLATFORM_AMD__ \
	    -o $@ $(HERE)ref_harness_st.c $(ST_SRC) $(SHIM_CF) $(HERE)../parallel_fem_codes_amd/compat/femgpu_compat.c \
	    -Wl,-rpath-link,/usr/lib/x86_64-linux-gnu -L$(LIBDIR) -Wl,-rpath,$(LIBDIR) -lfemgpu -L$(ROCM)/lib \
	    -Wl,-rpath,$(ROCM)/lib -lamdhip64 -L$(MPI)/lib -Wl,-rpath,$(MPI)/lib -lmpi -lm \
	    -Wl,--unresolved-symbols=ignore-in-object-files

$(OUT)/ns_harness: $(HERE)ref_harness_ns.c $(NS_SRC) $(NS_CF)
	mkdir -p $(OUT)
	gcc -O2 -fno-fast-math -ffp-contract=off -w -DSSNavierStokesEquations2D -I$(NS) -I$(MPI)/include \
	    -o $@ $(HERE)ref_harness_ns.c $(NS_SRC) $(NS_CF) \
	    -L$(MPI)/lib -Wl,-rpath,$(MPI)/lib -lmpi -lm -Wl,--unresolved-symbols=ignore-in-object-files

$(OUT)/eu_harness: $(HERE)ref_harness_eu.c $(EU_SRC) $(NS_CF)
	mkdir -p $(OUT)
	gcc -O2 -fno-fast-math -ffp-contract=off -w -DEulerEquations2D -I$(EU) -I$(MPI)/include \
	    -o $@ $(HERE)ref_harness_eu.c $(EU_SRC) $(NS_CF) \
	    -L$(MPI)/lib -Wl,-rpath,$(MPI)/lib -lmpi -lm -Wl,--unresolved-symbols=ignore-in-object-files

clean:
	rm -rf $(OUT)
