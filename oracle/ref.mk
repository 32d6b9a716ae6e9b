# Builds oracle/_ref/libpam_ref.so: routines of the reference's own Fortran
# sources, compiled unchanged from where they lie under $(REFSRC) with
# amdflang, plus the C-callable driver oracle/pam_ref_driver.f90.
# TEST INFRASTRUCTURE ONLY (pins of the C oracle); outputs go to oracle/_ref/.
#
# Only routines that need no library beyond the Fortran runtime are built.
# The whole passive path is not: radmat.f90:201-233 calls LAPACK, which this
# image does not have, and flang rejects fastemx.f90:287-288.
#
# -O2 follows the reference's default meson build type (debugoptimized,
# meson.build:7); the default x86-64 target has no FMA, so nothing is fused.

REF_FC ?= /opt/rocm/bin/amdflang
REFSRC ?= /root/reference/src
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
FFLAGS ?= -O2 -fPIC -w
# Only the driver's ref_* entry points are exported.  vars_atmosphere.f90:713
# calls GMTIME, a GNU extension flang's runtime lacks, from a routine the
# driver never calls: that one reference stays unresolved, and the library is
# loaded with lazy binding (oracle/refbind.py).
LDFLAGS_REF := -Wl,--version-script=$(HERE)ref_exports.map

# module dependency order
UNITS := kinds constants vars_index report_module nan rt_utilities settings conversions \
  descriptor_file vars_atmosphere drop_size_dist gasabs_module rosen98_gasabs tessem2 \
  e_sat_gg_water eps_ice eps_mix eps_water foam_utility_module fresnel liu \
  large_scale_correction_module small_scale_correction_module mie_scat_utilities \
  scat_utilities mie_spheres scatcnv rho_air make_mass_size make_dist_param make_dist \
  calc_moment check_print make_soft_spheroid
# the T-matrix producer: calc_tmatrix (tmatrix.f90) over Mishchenko's
# quad-precision EBCM code (tmatrix_amplq.lp.f, tmatrix_lpq.f).  zlib_stuff
# (module fileio) declares the zlib routines of the tmatrix_db = 'file' cache;
# the driver runs tmatrix_db = 'none', so they stay unresolved (lazy binding).
TM_F77 := tmatrix_lpq tmatrix_amplq.lp
TM_UNITS := zlib_stuff tmatrix

OBJS := $(addprefix $(OUT)/,$(addsuffix .o,$(UNITS) $(TM_F77) $(TM_UNITS)))

all: $(OUT)/libpam_ref.so

$(OUT)/libpam_ref.so: $(HERE)pam_ref_driver.f90 $(HERE)ref.mk $(HERE)ref_exports.map
	@mkdir -p $(OUT)
	@set -e; for u in $(UNITS); do \
	  $(REF_FC) $(FFLAGS) -module-dir $(OUT) -c $(REFSRC)/$$u.f90 -o $(OUT)/$$u.o; \
	done
	@set -e; for u in $(TM_F77); do \
	  $(REF_FC) $(FFLAGS) -I$(REFSRC) -module-dir $(OUT) -c $(REFSRC)/$$u.f -o $(OUT)/$$u.o; \
	done
	$(REF_FC) $(FFLAGS) -cpp -module-dir $(OUT) -c $(REFSRC)/zlib_stuff.f90 -o $(OUT)/zlib_stuff.o
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -c $(REFSRC)/tmatrix.f90 -o $(OUT)/tmatrix.o
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -c $(HERE)pam_ref_driver.f90 -o $(OUT)/pam_ref_driver.o
	$(REF_FC) -shared -fPIC $(LDFLAGS_REF) -o $@.tmp $(OBJS) $(OUT)/pam_ref_driver.o
	mv $@.tmp $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
