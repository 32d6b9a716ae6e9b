! C-callable driver around routines of the reference's own Fortran sources
! (igmk/pamtra src/*.f90), compiled unchanged by oracle/ref.mk into
! oracle/_ref/libpam_ref.so.
!
! TEST INFRASTRUCTURE ONLY.  The tests use it to pin the C oracle
! (oracle/pam_oracle.c) against the reference's compiled code, function by
! function.  Nothing in the product package loads it.  The passive path as a
! whole is not built here: its (I - R R)^-1 calls LAPACK (radmat.f90:201-233),
! which this image does not have, and fastemx.f90:287-288 is rejected by flang.
!
! Every routine takes plain C scalars and arrays.  Module state the
! reference routines read (settings flags, the drop_size_dist inputs, the
! layer's atmosphere) is set here before the call and the results copied out.

module pam_ref_driver
  use iso_c_binding
  use kinds, only: dbl, long
  implicit none

  interface
    subroutine rosen98_gasabs(errorstatus, freq, tempK, rhoWv, pres, absAir, absWv)
      import :: dbl, long
      integer(kind=long), intent(out) :: errorstatus
      real(kind=dbl), intent(in) :: freq, tempK, rhoWv, pres
      real(kind=dbl), intent(out) :: absAir, absWv
    end subroutine
    subroutine tessem2(errorstatus, freq, zenith_angle, wind_speed, temperature, salinity, &
                       emissivity, reflectivity)
      import :: dbl, long
      integer(kind=long), intent(out) :: errorstatus
      real(kind=dbl), intent(in) :: freq, zenith_angle, wind_speed, temperature, salinity
      real(kind=dbl), intent(out), dimension(2) :: emissivity, reflectivity
    end subroutine
    function e_sat_gg_water(T)
      import :: dbl
      real(kind=dbl) :: e_sat_gg_water
      real(kind=dbl), intent(in) :: T
    end function
    function eps_ice(T, f)
      import :: dbl
      complex(kind=dbl) :: eps_ice
      real(kind=dbl), intent(in) :: T, f
    end function
    function eps_mix(eps0, eps1, mix_de)
      import :: dbl
      complex(kind=dbl) :: eps_mix
      complex(kind=dbl), intent(in) :: eps0, eps1
      real(kind=dbl), intent(in) :: mix_de
    end function
    subroutine gausquad(n, xa, wt)
      import :: dbl
      integer :: n
      real(kind=dbl) :: xa(*), wt(*)
    end subroutine
    subroutine scatcnv(errorstatus, nlegen, coef, extinct, albedo, scatter_matrix, ext_matrix, &
                       emis_vec)
      integer*8 :: errorstatus
      integer :: nlegen
      real*8 :: coef(6, *), extinct, albedo
      real*8 :: scatter_matrix(*), ext_matrix(*), emis_vec(*)
    end subroutine
  end interface

contains

  ! rosen98_gasabs (rosen98_gasabs.f90:2-153): absorption by dry air and by
  ! water vapour, Np/km; freq GHz, T K, rho_v kg/m^3, p Pa.
  subroutine ref_gasabs(f, t, rhowv, p, absair, abswv, err) bind(C, name="ref_gasabs")
    real(c_double), value :: f, t, rhowv, p
    real(c_double), intent(out) :: absair, abswv
    integer(c_int), intent(out) :: err
    integer(kind=long) :: e
    real(kind=dbl) :: a, w
    call rosen98_gasabs(e, f, t, rhowv, p, a, w)
    absair = a
    abswv = w
    err = int(e, c_int)
  end subroutine

  ! tessem2 (tessem2.f90:1-137)
  subroutine ref_tessem2(f, zen, wind, temp, sal, emis, refl) bind(C, name="ref_tessem2")
    real(c_double), value :: f, zen, wind, temp, sal
    real(c_double), intent(out) :: emis(2), refl(2)
    integer(kind=long) :: e
    call tessem2(e, f, zen, wind, temp, sal, emis, refl)
  end subroutine

  ! The FASTEM5 pieces PAMTRA assembles in compute_fastemx (fastemx.f90:204-298):
  ! Liu permittivity, Fresnel reflectivities, foam coverage and reflectivity,
  ! large- and small-scale roughness corrections.
  subroutine ref_fastem_parts(f, zen, temp, sal, wind, out) bind(C, name="ref_fastem_parts")
    use liu, only: liu_ocean_permittivity
    use fresnel, only: fresnel_reflectivity
    use foam_utility_module, only: foam_coverage, foam_reflectivity
    use large_scale_correction_module, only: large_scale_correction
    use small_scale_correction_module, only: small_scale_correction
    use constants, only: deg2rad
    real(c_double), value :: f, zen, temp, sal, wind
    real(c_double), intent(out) :: out(11)
    complex(kind=dbl) :: eps
    real(kind=dbl) :: cz, rv, rh, cover, rvf, rhf, rvl, rhl, corr
    call liu_ocean_permittivity(temp, sal, f, eps)
    cz = cos(zen * deg2rad)
    call fresnel_reflectivity(eps, cz, rv, rh)
    call foam_coverage(wind, cover)
    call foam_reflectivity(zen, f, rvf, rhf)
    call large_scale_correction(f, cz, wind, rvl, rhl)
    call small_scale_correction(f, cz, wind, corr)
    out = [real(eps, dbl), aimag(eps), rv, rh, cover, rvf, rhf, rvl, rhl, corr, cz]
  end subroutine

  ! eps_water_tkc (eps_water.f90:653-757): T degC, f GHz
  subroutine ref_eps_water_tkc(t, f, re, im) bind(C, name="ref_eps_water_tkc")
    use eps_water, only: eps_water_tkc
    real(c_double), value :: t, f
    real(c_double), intent(out) :: re, im
    complex(kind=dbl) :: eps
    call eps_water_tkc(t, f, eps)
    re = real(eps, dbl)
    im = aimag(eps)
  end subroutine

  ! eps_ice (eps_ice.f90:1-55): T K, f GHz
  subroutine ref_eps_ice(t, f, re, im) bind(C, name="ref_eps_ice")
    real(c_double), value :: t, f
    real(c_double), intent(out) :: re, im
    complex(kind=dbl) :: eps
    eps = eps_ice(t, f)
    re = real(eps, dbl)
    im = aimag(eps)
  end subroutine

  ! eps_mix (eps_mix.f90:1-48) of (1, 0) and eps1, as mie_spheres.f90:139-144 calls it
  subroutine ref_eps_mix(re1, im1, mix_de, re, im) bind(C, name="ref_eps_mix")
    real(c_double), value :: re1, im1, mix_de
    real(c_double), intent(out) :: re, im
    complex(kind=dbl) :: e
    e = eps_mix((1.d0, 0.d0), cmplx(re1, im1, kind=dbl), mix_de)
    re = real(e, dbl)
    im = aimag(e)
  end subroutine

  ! e_sat_gg_water (e_sat_gg_water.f90:1-21), hPa
  function ref_esat(t) bind(C, name="ref_esat") result(e)
    real(c_double), value :: t
    real(c_double) :: e
    e = e_sat_gg_water(t)
  end function

  ! q2abs (conversions.f90:5-24)
  function ref_q2abs(spec, t, p, qv, qall) bind(C, name="ref_q2abs") result(r)
    use conversions, only: q2abs
    real(c_double), value :: spec, t, p, qv, qall
    real(c_double) :: r
    r = q2abs(spec, t, p, qv, qall)
  end function

  ! LOBATTO_QUADRATURE (rt_utilities.f90:165-203) and gausquad (scat_utilities.f90:1-44)
  subroutine ref_lobatto(n, mu, w) bind(C, name="ref_lobatto")
    use rt_utilities, only: lobatto_quadrature
    integer(c_int), value :: n
    real(c_double), intent(out) :: mu(n), w(n)
    call lobatto_quadrature(int(n), mu, w)
  end subroutine

  subroutine ref_gausquad(n, x, w) bind(C, name="ref_gausquad")
    integer(c_int), value :: n
    real(c_double), intent(out) :: x(n), w(n)
    call gausquad(int(n), x, w)
  end subroutine

  ! miecalc + miecross (mie_scat_utilities.f90:8-116): the Mie coefficients
  ! a_n, b_n (n = 1..nterms, the stop index of x) and the efficiencies.
  ! maxn: room in the output arrays.  Returns the error status.
  function ref_mie(x, mre, mim, maxn, nterms, a, b, q) bind(C, name="ref_mie") result(st)
    use mie_scat_utilities, only: miecalc, miecross
    real(c_double), value :: x, mre, mim
    integer(c_int), value :: maxn
    integer(c_int), intent(out) :: nterms
    real(c_double), intent(out) :: a(2, maxn), b(2, maxn), q(3)
    integer(c_int) :: st
    complex(kind=dbl), allocatable :: ca(:), cb(:)
    integer(kind=long) :: e
    integer :: nt, i
    allocate(ca(10015), cb(10015))
    nt = 0
    call miecalc(e, nt, x, cmplx(mre, -mim, kind=dbl), ca, cb)
    st = int(e, c_int)
    nterms = nt
    if (e /= 0) return
    call miecross(nt, x, ca, cb, q(1), q(2), q(3))
    do i = 1, min(nt, int(maxn))
      a(1, i) = real(ca(i), dbl)
      a(2, i) = aimag(ca(i))
      b(1, i) = real(cb(i), dbl)
      b(2, i) = aimag(cb(i))
    end do
  end function

  ! mieangle (mie_scat_utilities.f90:119-160) at mu for the sphere's own
  ! series: p(1:4) = p1..p4
  subroutine ref_mie_angle(x, mre, mim, mu, p) bind(C, name="ref_mie_angle")
    use mie_scat_utilities, only: miecalc, mieangle
    real(c_double), value :: x, mre, mim, mu
    real(c_double), intent(out) :: p(4)
    complex(kind=dbl), allocatable :: ca(:), cb(:)
    integer(kind=long) :: e
    integer :: nt
    allocate(ca(10015), cb(10015))
    nt = 0
    call miecalc(e, nt, x, cmplx(mre, -mim, kind=dbl), ca, cb)
    call mieangle(nt, ca, cb, mu, p(1), p(2), p(3), p(4))
  end subroutine

  ! calc_mie_spheres (mie_spheres.f90:14-345) with lhyd_absorption on:
  ! out = (extinction, albedo, back_scatt); legen(200, 4) = legen..legen4
  function ref_mie_spheres(freq, t, liq_ice, nbins, diameter, del_d, ndens, density, refre, &
                           refim, out, legen, back_spec, nlegen) &
      bind(C, name="ref_mie_spheres") result(st)
    use mie_spheres, only: calc_mie_spheres
    use settings, only: lhyd_absorption, maxnleg
    real(c_double), value :: freq, t, refre, refim
    integer(c_int), value :: liq_ice, nbins
    real(c_double), intent(in) :: diameter(nbins), del_d(nbins), ndens(nbins), density(nbins)
    real(c_double), intent(out) :: out(3), legen(maxnleg, 4), back_spec(nbins)
    integer(c_int), intent(out) :: nlegen
    integer(c_int) :: st
    integer(kind=long) :: e
    integer :: nl
    lhyd_absorption = .true.
    call calc_mie_spheres(e, freq, t, int(liq_ice), int(nbins), diameter, del_d, ndens, density, &
                          refre, refim, out(1), out(2), out(3), nl, legen(:, 1), legen(:, 2), &
                          legen(:, 3), legen(:, 4), back_spec)
    nlegen = nl
    st = int(e, c_int)
  end function

  ! scatcnv (scatcnv.f90:1-192): coef(6, 200) with rows P11, P12, P33, P34,
  ! P11, P33 as calc_scatProperties fills them (scatProperties.f90:930-937).
  ! smat: scatter_matrix(nstokes, nummu, nstokes, nummu, 4) in Fortran order.
  subroutine ref_scatcnv(nlegen, coef, kext, salb, smat, ext, emis) bind(C, name="ref_scatcnv")
    integer(c_int), value :: nlegen
    real(c_double), intent(in) :: coef(6, 200)
    real(c_double), value :: kext, salb
    real(c_double), intent(out) :: smat(*), ext(*), emis(*)
    integer*8 :: e
    real*8 :: c(6, 200)
    c = coef
    e = 0
    call scatcnv(e, int(nlegen), c, kext, salb, smat, ext, emis)
  end subroutine

  ! calc_tmatrix (tmatrix.f90:14-444) for one hydrometeor of one layer, passive
  ! only (active = .false.: every incidence angle), tmatrix_db = 'none', one
  ! radar polarisation 'NN'.  freq in Hz; (mre, mim) the refractive index
  ! m = mre - i mim the reference's scatProperties.f90:222 passes.
  ! smat(2,16,2,16,2), ext(2,2,16), emis(2,16) as calc_tmatrix returns them;
  ! sback(2,2) starts at zero (calc_tmatrix only adds to it).
  function ref_calc_tmatrix(freq, mre, mim, phase, nbins, dmax, del_d, ndens, density, &
                            as_ratio, canting, temp, smat, ext, emis, sback) &
                            bind(C, name="ref_calc_tmatrix") result(st)
    use tmatrix, only: calc_tmatrix
    use settings, only: active, passive, tmatrix_db, radar_npol, radar_pol, &
                        lhyd_absorption, nummu, nstokes
    real(c_double), value :: freq, mre, mim, temp
    integer(c_int), value :: phase, nbins
    real(c_double), intent(in) :: dmax(nbins), del_d(nbins), ndens(nbins), density(nbins), &
                                  as_ratio(nbins), canting(nbins)
    real(c_double), intent(out) :: smat(nstokes, nummu, nstokes, nummu, 2), &
                                   ext(nstokes, nstokes, nummu), emis(nstokes, nummu), sback(2, 2)
    integer(c_int) :: st
    integer(kind=long) :: e
    real(kind=dbl) :: back_spec(1, nbins)
    active = .false.
    passive = .true.
    tmatrix_db = 'none'
    radar_npol = 1
    radar_pol(1) = 'NN'
    lhyd_absorption = .true.
    sback = 0.d0
    e = 0
    call calc_tmatrix(e, freq, cmplx(mre, -mim, kind=dbl), int(phase), int(nbins), dmax, del_d, &
                      ndens, density, as_ratio, canting, temp, smat, ext, emis, back_spec, sback)
    st = int(e, c_int)
  end function

  ! tmatrix_amplq (tmatrix_amplq.lp.f:276-540) for one spheroid (NP = -1,
  ! RAT = 1): nmax and the REAL*4 T matrix of COMMON /TMAT/ as
  ! t(n', n, m, c) = t[c][m][n-1][n'-1] with the strides (cap+1, cap, cap) of a
  ! C array, c = RT11, RT12, RT21, RT22, IT11, IT12, IT21, IT22.
  function ref_tmatrix_amplq(lam, mrr, mri, axi, as_ratio, t, cap) &
                             bind(C, name="ref_tmatrix_amplq") result(nmax)
    real(c_double), value :: lam, mrr, mri, axi, as_ratio
    integer(c_int), value :: cap
    real(c_float), intent(out) :: t(cap, cap, 0:cap, 8)
    integer(c_int) :: nmax
    real*4 :: rt11(101, 100, 100), rt12(101, 100, 100), rt21(101, 100, 100), &
              rt22(101, 100, 100), it11(101, 100, 100), it12(101, 100, 100), &
              it21(101, 100, 100), it22(101, 100, 100)
    common /tmat/ rt11, rt12, rt21, rt22, it11, it12, it21, it22
    real*16 :: ql, qr, qi
    real*8 :: ax, ar, rat
    integer :: np, nm, m, n, nn
    integer(kind=long) :: e
    interface
      subroutine tmatrix_amplq(errorstatus, lam, mrr, mri, axi, as_ratio, rat, np, nmax)
        import :: long
        integer(kind=long) :: errorstatus
        real*16 :: lam, mrr, mri
        real*8 :: axi, as_ratio, rat
        integer :: np, nmax
      end subroutine
    end interface
    ql = lam
    qr = mrr
    qi = mri
    ax = axi
    ar = as_ratio
    rat = 1.d0
    np = -1
    e = 0
    call tmatrix_amplq(e, ql, qr, qi, ax, ar, rat, np, nm)
    nmax = 0
    if (e /= 0) return
    nmax = nm
    if (nm > cap) return
    do m = 0, nm
      do n = 1, nm
        do nn = 1, nm
          t(nn, n, m, 1) = rt11(m + 1, n, nn)
          t(nn, n, m, 2) = rt12(m + 1, n, nn)
          t(nn, n, m, 3) = rt21(m + 1, n, nn)
          t(nn, n, m, 4) = rt22(m + 1, n, nn)
          t(nn, n, m, 5) = it11(m + 1, n, nn)
          t(nn, n, m, 6) = it12(m + 1, n, nn)
          t(nn, n, m, 7) = it21(m + 1, n, nn)
          t(nn, n, m, 8) = it22(m + 1, n, nn)
        end do
      end do
    end do
  end function

  ! run_drop_size_dist (drop_size_dist.f90:107-218) for one (layer,
  ! hydrometeor), with the module inputs hydrometeor_extinction.f90:150-300
  ! sets.  dreal(1:14) = as_ratio, rho_ms, a_ms, b_ms, alpha_as, beta_as,
  ! p_1, p_2, p_3, p_4, d_1, d_2, canting, softsphere_min_density;
  ! dint(1:7) = liq_ice, moment_in, nbin, adaptive grid, conserve mass,
  ! limit density/area, sfc_type; layer(1:5) = q_h, n_tot, r_eff, T, p.
  ! Outputs nbin values each: d_ds, delta_d_ds, n_ds, density2scat,
  ! diameter2scat, mass_ds, area_ds.  Returns the error status.
  function ref_dsd(dist, dreal, dint, layer, nbin_out, d_ds_o, delta_o, n_ds_o, rho_o, dscat_o, &
                   mass_o, area_o) bind(C, name="ref_dsd") result(st)
    use drop_size_dist
    use settings, only: hydro_adaptive_grid, conserve_mass_rescale_dsd, &
                        hydro_limit_density_area, hydro_softsphere_min_density
    use vars_index, only: i_x, i_y, i_z, i_h, i_f
    use vars_atmosphere, only: sfc_type, atmo_press, atmo_temp
    character(kind=c_char), intent(in) :: dist(16)
    real(c_double), intent(in) :: dreal(14), layer(5)
    integer(c_int), intent(in) :: dint(7)
    integer(c_int), intent(out) :: nbin_out
    real(c_double), intent(out) :: d_ds_o(*), delta_o(*), n_ds_o(*), rho_o(*), dscat_o(*), &
                                   mass_o(*), area_o(*)
    integer(c_int) :: st
    integer(kind=long) :: e
    integer :: k
    character(len=15) :: name
    name = ' '
    do k = 1, 15
      if (dist(k) == c_null_char) exit
      name(k:k) = dist(k)
    end do
    i_x = 1; i_y = 1; i_z = 1; i_h = 1; i_f = 1
    if (.not. allocated(sfc_type)) allocate(sfc_type(1, 1))
    if (.not. allocated(atmo_press)) allocate(atmo_press(1, 1, 1))
    if (.not. allocated(atmo_temp)) allocate(atmo_temp(1, 1, 1))
    sfc_type(1, 1) = dint(7)
    atmo_temp(1, 1, 1) = layer(4)
    atmo_press(1, 1, 1) = layer(5)
    hydro_name = 'h'
    dist_name = name
    as_ratio = dreal(1); rho_ms = dreal(2); a_ms = dreal(3); b_ms = dreal(4)
    alpha_as = dreal(5); beta_as = dreal(6)
    p_1 = dreal(7); p_2 = dreal(8); p_3 = dreal(9); p_4 = dreal(10)
    d_1 = dreal(11); d_2 = dreal(12); dsd_canting = dreal(13)
    hydro_softsphere_min_density = dreal(14)
    liq_ice = dint(1); moment_in = dint(2); nbin = dint(3)
    if (trim(dist_name) == 'mono' .or. trim(dist_name) == 'mono_cosmo_ice') nbin = 2
    hydro_adaptive_grid = dint(4) /= 0
    conserve_mass_rescale_dsd = dint(5) /= 0
    hydro_limit_density_area = dint(6) /= 0
    q_h = layer(1); n_tot = layer(2); r_eff = layer(3); layer_t = layer(4); pressure = layer(5)
    call deallocateVars_drop_size_dist()
    call allocateVars_drop_size_dist()
    e = 0
    call run_drop_size_dist(e)
    st = int(e, c_int)
    nbin_out = nbin
    if (e == 2) return
    d_ds_o(1:nbin) = d_ds
    delta_o(1:nbin) = delta_d_ds
    n_ds_o(1:nbin) = n_ds
    rho_o(1:nbin) = density2scat
    dscat_o(1:nbin) = diameter2scat
    mass_o(1:nbin) = mass_ds
    area_o(1:nbin) = area_ds
  end function

end module pam_ref_driver
