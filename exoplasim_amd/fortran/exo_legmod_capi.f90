! exo_legmod_capi.f90 -- ISO_C_BINDING interfaces of libexosht.so
! (include/exo_sht.h) for the Fortran drop-in subroutines.
module exo_legmod_capi
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function exo_inigau(klat, pz0, pzw) bind(C, name="exo_inigau")
      import
      integer(c_int), value :: klat
      type(c_ptr), value :: pz0, pzw
    end function
    integer(c_int) function exo_legini(ntru, nlat, nlev, sid, gwd, skgpsp, skspgp, plavor) &
        bind(C, name="exo_legini")
      import
      integer(c_int), value :: ntru, nlat, nlev
      type(c_ptr), value :: sid, gwd, skgpsp, skspgp
      real(c_double), value :: plavor
    end function
  end interface
  ! the entry points with (ptr, ptr, int) and (ptr x4) / (ptr x9) shapes
  interface
    integer(c_int) function exo_sp2fc_f64(x, y, n) bind(C, name="exo_sp2fc_f64")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_sp2fc_f32(x, y, n) bind(C, name="exo_sp2fc_f32")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_sp2fcdmu_f64(x, y, n) bind(C, name="exo_sp2fcdmu_f64")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_sp2fcdmu_f32(x, y, n) bind(C, name="exo_sp2fcdmu_f32")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_fc2sp_f64(x, y, n) bind(C, name="exo_fc2sp_f64")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_fc2sp_f32(x, y, n) bind(C, name="exo_fc2sp_f32")
      import
      type(c_ptr), value :: x, y
      integer(c_int), value :: n
    end function
    integer(c_int) function exo_gp2fc_f64(a, n, lot) bind(C, name="exo_gp2fc_f64")
      import
      type(c_ptr), value :: a
      integer(c_int), value :: n, lot
    end function
    integer(c_int) function exo_gp2fc_f32(a, n, lot) bind(C, name="exo_gp2fc_f32")
      import
      type(c_ptr), value :: a
      integer(c_int), value :: n, lot
    end function
    integer(c_int) function exo_fc2gp_f64(a, n, lot) bind(C, name="exo_fc2gp_f64")
      import
      type(c_ptr), value :: a
      integer(c_int), value :: n, lot
    end function
    integer(c_int) function exo_fc2gp_f32(a, n, lot) bind(C, name="exo_fc2gp_f32")
      import
      type(c_ptr), value :: a
      integer(c_int), value :: n, lot
    end function
    integer(c_int) function exo_dv2uv_f64(a, b, c, d) bind(C, name="exo_dv2uv_f64")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_dv2uv_f32(a, b, c, d) bind(C, name="exo_dv2uv_f32")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_uv2dv_f64(a, b, c, d) bind(C, name="exo_uv2dv_f64")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_uv2dv_f32(a, b, c, d) bind(C, name="exo_uv2dv_f32")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_qtend_f64(a, b, c, d) bind(C, name="exo_qtend_f64")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_qtend_f32(a, b, c, d) bind(C, name="exo_qtend_f32")
      import
      type(c_ptr), value :: a, b, c, d
    end function
    integer(c_int) function exo_mktend_f64(a, b, c, d, e, f, g, h, i) bind(C, name="exo_mktend_f64")
      import
      type(c_ptr), value :: a, b, c, d, e, f, g, h, i
    end function
    integer(c_int) function exo_mktend_f32(a, b, c, d, e, f, g, h, i) bind(C, name="exo_mktend_f32")
      import
      type(c_ptr), value :: a, b, c, d, e, f, g, h, i
    end function
  end interface
contains
  ! a failed transform ends the run through the model's abort path
  subroutine exo_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(len=80) :: msg
    if (rc /= 0) then
      write(msg, '(a,a,a,i0)') 'exoplasim_amd: ', what, ' failed with code ', rc
      call exo_abort(msg)
    endif
  end subroutine
end module exo_legmod_capi
