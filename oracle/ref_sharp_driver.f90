! ref_sharp_driver.f90 -- TEST INFRASTRUCTURE: sets the SharpClaw module state the
! way SharpClawSolver._set_fortran_parameters does (sharpclaw/solver.py:563-588),
! so the reference's compiled flux2/flux1 can be called from C.  Built only into
! oracle/_ref/ by oracle/ref.mk together with the reference's own sources.
subroutine ref_sharp_init(num_dim_, lim_type_, weno_order_, char_decomp_, num_waves_, &
                          dxs, mthlim_, maxnx, num_ghost, num_eqn) bind(C, name="ref_sharp_init")
    use iso_c_binding
    use ClawParams
    use workspace
    use reconstruct
    implicit none
    integer(c_int), value :: num_dim_, lim_type_, weno_order_, char_decomp_, num_waves_
    integer(c_int), value :: maxnx, num_ghost, num_eqn
    real(c_double), intent(in) :: dxs(*)
    integer(c_int), intent(in) :: mthlim_(*)

    if (work_alloc) call dealloc_workspace(char_decomp)
    if (recon_alloc) then
        call dealloc_recon_workspace(lim_type, char_decomp)
        recon_alloc = .false.
    end if
    call dealloc_clawparams()

    num_dim = num_dim_
    lim_type = lim_type_
    weno_order = weno_order_
    char_decomp = char_decomp_
    tfluct_solver = .false.
    fwave = .false.
    index_capa = 0
    num_waves = num_waves_
    call alloc_clawparams()
    xlower = 0.d0
    xupper = 1.d0
    dx(1:num_dim) = dxs(1:num_dim)
    mthlim(1:num_waves) = mthlim_(1:num_waves)
    call alloc_workspace(maxnx, num_ghost, num_eqn, num_waves, char_decomp)
    call alloc_recon_workspace(maxnx, num_ghost, num_eqn, num_waves, lim_type, char_decomp)
end subroutine ref_sharp_init
