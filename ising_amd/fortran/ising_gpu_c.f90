! ising_gpu_c.f90 -- Fortran bind(C) view of include/ising_gpu.h.
!
! Derived types mirror the C structures field for field; every C entry point
! is declared with its exact argument passing (value scalars, c_ptr handles,
! arrays by reference). The uint64 seed is passed as a 64-bit integer with
! the same bits.
module ising_gpu_c
        use iso_c_binding
        implicit none

        integer(c_int), parameter :: ISING_OK = 0
        integer(c_int), parameter :: ISING_ERR_INVALID_ARG = -1
        integer(c_int), parameter :: ISING_ERR_BAD_DIMS = -2
        integer(c_int), parameter :: ISING_ERR_BAD_SPIN = -3
        integer(c_int), parameter :: ISING_ERR_NO_DEVICE = -4
        integer(c_int), parameter :: ISING_ERR_HIP = -5
        integer(c_int), parameter :: ISING_ERR_UNSUPPORTED = -6
        integer(c_int), parameter :: ISING_ERR_OUT_OF_MEMORY = -7
        integer(c_int), parameter :: ISING_ERR_STATE = -8

        integer(c_int32_t), parameter :: ISING_DMI_NONE = 0
        integer(c_int32_t), parameter :: ISING_DMI_BULK = 1
        integer(c_int32_t), parameter :: ISING_DMI_INTERFACIAL_REF = 2

        ! launch tuning flags (ISING_TUNE_FUSED(lc) = lc * 2**ISING_TUNE_FUSED_SHIFT,
        ! ISING_TUNE_SLAB(hy) = hy * 2**ISING_TUNE_SLAB_SHIFT)
        integer(c_int32_t), parameter :: ISING_TUNE_XCD_REMAP = 1
        integer(c_int32_t), parameter :: ISING_TUNE_NT_OWN = 2
        integer(c_int32_t), parameter :: ISING_TUNE_FUSED_SHIFT = 8
        integer(c_int32_t), parameter :: ISING_TUNE_SLAB_SHIFT = 16
        integer(c_int), parameter :: ISING_TUNE_MAX_CANDIDATES = 16
        integer(c_int), parameter :: ISING_MLP_PARAMS = 1793
        integer(c_int), parameter :: ISING_MLP_RECORD = 32

        ! the reference's constants::mu_0 (src/constants.f90:15), single-precision literal
        real(c_double), parameter :: ISING_MU0_REF = 1.256637006008532e-06_c_double

        integer(c_int32_t), parameter :: ISING_CELL_SC = 0
        integer(c_int32_t), parameter :: ISING_CELL_BCC = 1
        integer(c_int32_t), parameter :: ISING_CELL_HYPERKAGOME = 2

        type, bind(C) :: ising_lattice
                integer(c_int32_t) :: lx, ly, lz
                integer(c_int32_t) :: periodic(3)
                integer(c_int32_t) :: cell = 0
        end type ising_lattice

        type, bind(C) :: ising_hamiltonian
                real(c_double) :: J, D
                integer(c_int32_t) :: dmi_type, demag
                real(c_double) :: Bz, Kz, zeeman_total_scale
        end type ising_hamiltonian

        type, bind(C) :: ising_ladder_desc
                type(ising_lattice) :: lattice
                type(ising_hamiltonian) :: ham
                integer(c_int32_t) :: n_replicas, reserved
                real(c_double) :: cone_radius
                type(c_ptr) :: betas
        end type ising_ladder_desc

        type, bind(C) :: ising_schedule
                integer(c_int32_t) :: n_sweeps, swap_every, measure_every, winding_every, use_graph
                integer(c_int32_t) :: winding_layer   ! 0: every layer; z+1: layer z only, series [n_q][R]
        end type ising_schedule

        type, bind(C) :: ising_series
                type(c_ptr) :: energy = c_null_ptr
                type(c_ptr) :: magnetisation = c_null_ptr
                type(c_ptr) :: winding = c_null_ptr
                type(c_ptr) :: winding_raw = c_null_ptr
        end type ising_series

        type, bind(C) :: ising_skyrmion_params
                integer(c_int32_t) :: scale_x = 10, scale_y = 10
                integer(c_int32_t) :: n_particles = 1
                integer(c_int32_t) :: num_thresholds = 60
                integer(c_int32_t) :: num_sigmas = 20
                integer(c_int32_t) :: reserved = 0
                real(c_double) :: min_threshold = 1.0e-5_c_double
                real(c_double) :: max_threshold = 0.95_c_double
        end type ising_skyrmion_params

        interface
                function ising_version() bind(C, name="ising_version") result(p)
                        import :: c_ptr
                        type(c_ptr) :: p
                end function
                function ising_device_count(n) bind(C, name="ising_device_count") result(rc)
                        import :: c_int
                        integer(c_int), intent(out) :: n
                        integer(c_int) :: rc
                end function
                function ising_create(ctx, ladders, n_ladders, seed, device) bind(C, name="ising_create") result(rc)
                        import :: c_ptr, c_int, c_int64_t, ising_ladder_desc
                        type(c_ptr), intent(out) :: ctx
                        type(ising_ladder_desc), intent(in) :: ladders(*)
                        integer(c_int), value :: n_ladders
                        integer(c_int64_t), value :: seed
                        integer(c_int), value :: device
                        integer(c_int) :: rc
                end function
                subroutine ising_destroy(ctx) bind(C, name="ising_destroy")
                        import :: c_ptr
                        type(c_ptr), value :: ctx
                end subroutine
                function ising_last_error(ctx) bind(C, name="ising_last_error") result(p)
                        import :: c_ptr
                        type(c_ptr), value :: ctx
                        type(c_ptr) :: p
                end function
                function ising_num_replicas(ctx, n) bind(C, name="ising_num_replicas") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), intent(out) :: n
                        integer(c_int) :: rc
                end function
                function ising_replica_sites(ctx, replica, n) bind(C, name="ising_replica_sites") result(rc)
                        import :: c_ptr, c_int, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        integer(c_int64_t), intent(out) :: n
                        integer(c_int) :: rc
                end function
                function ising_set_hamiltonian(ctx, ladder, h) bind(C, name="ising_set_hamiltonian") result(rc)
                        import :: c_ptr, c_int, ising_hamiltonian
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: ladder
                        type(ising_hamiltonian), intent(in) :: h
                        integer(c_int) :: rc
                end function
                function ising_set_cone(ctx, ladder, radius) bind(C, name="ising_set_cone") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: ladder
                        real(c_double), value :: radius
                        integer(c_int) :: rc
                end function
                function ising_set_beta(ctx, replica, beta) bind(C, name="ising_set_beta") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), value :: beta
                        integer(c_int) :: rc
                end function
                function ising_get_betas(ctx, betas) bind(C, name="ising_get_betas") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        real(c_double), intent(out) :: betas(*)
                        integer(c_int) :: rc
                end function
                function ising_set_spins(ctx, replica, soa, n) bind(C, name="ising_set_spins") result(rc)
                        import :: c_ptr, c_int, c_double, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), intent(in) :: soa(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_get_spins(ctx, replica, soa, n) bind(C, name="ising_get_spins") result(rc)
                        import :: c_ptr, c_int, c_double, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), intent(out) :: soa(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_init_random(ctx, replica) bind(C, name="ising_init_random") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        integer(c_int) :: rc
                end function
                function ising_init_ferro(ctx, replica, frac) bind(C, name="ising_init_ferro") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), value :: frac
                        integer(c_int) :: rc
                end function
                function ising_sweep(ctx, n_sweeps) bind(C, name="ising_sweep") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: n_sweeps
                        integer(c_int) :: rc
                end function
                function ising_energy(ctx, e) bind(C, name="ising_energy") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        real(c_double), intent(out) :: e(*)
                        integer(c_int) :: rc
                end function
                function ising_energy_with(ctx, replica, h, e) bind(C, name="ising_energy_with") result(rc)
                        import :: c_ptr, c_int, c_double, ising_hamiltonian
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        type(ising_hamiltonian), intent(in) :: h
                        real(c_double), intent(out) :: e
                        integer(c_int) :: rc
                end function
                function ising_magnetisation(ctx, m) bind(C, name="ising_magnetisation") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        real(c_double), intent(out) :: m(*)
                        integer(c_int) :: rc
                end function
                function ising_winding(ctx, replica, z, q, q_raw) bind(C, name="ising_winding") result(rc)
                        import :: c_ptr, c_int, c_double, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica, z
                        integer(c_int64_t), intent(out) :: q
                        real(c_double), intent(out) :: q_raw
                        integer(c_int) :: rc
                end function
                function ising_winding_density(ctx, replica, z, density) bind(C, name="ising_winding_density") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica, z
                        real(c_double), intent(out) :: density(*)
                        integer(c_int) :: rc
                end function
                function ising_winding_density_interpolated(ctx, replica, z, scale_x, scale_y, density) &
                                bind(C, name="ising_winding_density_interpolated") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica, z, scale_x, scale_y
                        real(c_double), intent(out) :: density(*)
                        integer(c_int) :: rc
                end function
                function ising_skyrmion_distribution(ctx, replica, z, params, counts, table) &
                                bind(C, name="ising_skyrmion_distribution") result(rc)
                        import :: c_ptr, c_int, c_int32_t, ising_skyrmion_params
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica, z
                        type(ising_skyrmion_params), intent(in) :: params
                        integer(c_int32_t), intent(out) :: counts(*)
                        type(c_ptr), value :: table
                        integer(c_int) :: rc
                end function
                function ising_winding_all(ctx, q, q_raw) bind(C, name="ising_winding_all") result(rc)
                        import :: c_ptr, c_int, c_double, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int64_t), intent(out) :: q(*)
                        real(c_double), intent(out) :: q_raw(*)
                        integer(c_int) :: rc
                end function
                function ising_delta_energy(ctx, replica, sites, props, n, de) bind(C, name="ising_delta_energy") result(rc)
                        import :: c_ptr, c_int, c_int32_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        integer(c_int32_t), intent(in) :: sites(*)
                        real(c_double), intent(in) :: props(*)
                        integer(c_int), value :: n
                        real(c_double), intent(out) :: de(*)
                        integer(c_int) :: rc
                end function
                function ising_pt_run(ctx, sched, series) bind(C, name="ising_pt_run") result(rc)
                        import :: c_ptr, c_int, ising_schedule, ising_series
                        type(c_ptr), value :: ctx
                        type(ising_schedule), intent(in) :: sched
                        type(ising_series), intent(in) :: series
                        integer(c_int) :: rc
                end function
                function ising_pt_run_static(ctx, sched, series) bind(C, name="ising_pt_run_static") result(rc)
                        import :: c_ptr, c_int, ising_schedule, ising_series
                        type(c_ptr), value :: ctx
                        type(ising_schedule), intent(in) :: sched
                        type(ising_series), intent(in) :: series
                        integer(c_int) :: rc
                end function
                function ising_save_state(ctx) bind(C, name="ising_save_state") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int) :: rc
                end function
                function ising_restore_state(ctx) bind(C, name="ising_restore_state") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int) :: rc
                end function
                function ising_temperature_map(ctx, slot_of_temp) bind(C, name="ising_temperature_map") result(rc)
                        import :: c_ptr, c_int, c_int32_t
                        type(c_ptr), value :: ctx
                        integer(c_int32_t), intent(out) :: slot_of_temp(*)
                        integer(c_int) :: rc
                end function
                function ising_swap_stats(ctx, attempts, accepts) bind(C, name="ising_swap_stats") result(rc)
                        import :: c_ptr, c_int, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int64_t), intent(out) :: attempts(*), accepts(*)
                        integer(c_int) :: rc
                end function
                function ising_accept_counts(ctx, accepted) bind(C, name="ising_accept_counts") result(rc)
                        import :: c_ptr, c_int, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int64_t), intent(out) :: accepted(*)
                        integer(c_int) :: rc
                end function
                function ising_synchronize(ctx) bind(C, name="ising_synchronize") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int) :: rc
                end function
                function ising_time_sweeps(ctx, n_sweeps, ms_per_launch) bind(C, name="ising_time_sweeps") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: n_sweeps
                        real(c_double), intent(out) :: ms_per_launch
                        integer(c_int) :: rc
                end function
                function ising_bandwidth_probe(ctx, bytes, ms) bind(C, name="ising_bandwidth_probe") result(rc)
                        import :: c_ptr, c_int, c_int64_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int64_t), value :: bytes
                        real(c_double), intent(out) :: ms
                        integer(c_int) :: rc
                end function
                function ising_set_tuning(ctx, flags) bind(C, name="ising_set_tuning") result(rc)
                        import :: c_ptr, c_int, c_int32_t
                        type(c_ptr), value :: ctx
                        integer(c_int32_t), value :: flags
                        integer(c_int) :: rc
                end function
                function ising_get_tuning(ctx, flags) bind(C, name="ising_get_tuning") result(rc)
                        import :: c_ptr, c_int, c_int32_t
                        type(c_ptr), value :: ctx
                        integer(c_int32_t), intent(out) :: flags
                        integer(c_int) :: rc
                end function
                function ising_tuning_candidates(ctx, flags, count) bind(C, name="ising_tuning_candidates") result(rc)
                        import :: c_ptr, c_int, c_int32_t
                        type(c_ptr), value :: ctx
                        integer(c_int32_t), intent(out) :: flags(*)
                        integer(c_int), intent(out) :: count
                        integer(c_int) :: rc
                end function
                function ising_autotune(ctx, n_sweeps, best, ms) bind(C, name="ising_autotune") result(rc)
                        import :: c_ptr, c_int, c_int32_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: n_sweeps
                        integer(c_int32_t), intent(out) :: best
                        real(c_double), intent(out) :: ms(*)
                        integer(c_int) :: rc
                end function
                function ising_mlp_set_params(ctx, params, n) bind(C, name="ising_mlp_set_params") result(rc)
                        import :: c_ptr, c_int, c_float
                        type(c_ptr), value :: ctx
                        real(c_float), intent(in) :: params(*)
                        integer(c_int), value :: n
                        integer(c_int) :: rc
                end function
                function ising_mlp_get_params(ctx, params, n) bind(C, name="ising_mlp_get_params") result(rc)
                        import :: c_ptr, c_int, c_float
                        type(c_ptr), value :: ctx
                        real(c_float), intent(out) :: params(*)
                        integer(c_int), value :: n
                        integer(c_int) :: rc
                end function
                function ising_mlp_enable(ctx, ladder, on) bind(C, name="ising_mlp_enable") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: ladder, on
                        integer(c_int) :: rc
                end function
                function ising_mlp_collect(ctx, per_replica) bind(C, name="ising_mlp_collect") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: per_replica
                        integer(c_int) :: rc
                end function
                function ising_mlp_num_records(ctx, n) bind(C, name="ising_mlp_num_records") result(rc)
                        import :: c_ptr, c_int, c_int64_t
                        type(c_ptr), value :: ctx
                        integer(c_int64_t), intent(out) :: n
                        integer(c_int) :: rc
                end function
                function ising_mlp_clear_records(ctx) bind(C, name="ising_mlp_clear_records") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int) :: rc
                end function
                function ising_mlp_get_records(ctx, out, n) bind(C, name="ising_mlp_get_records") result(rc)
                        import :: c_ptr, c_int, c_int64_t, c_float
                        type(c_ptr), value :: ctx
                        real(c_float), intent(out) :: out(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_mlp_train(ctx, steps, batch, lr, loss) bind(C, name="ising_mlp_train") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: steps, batch
                        real(c_double), value :: lr
                        real(c_double), intent(out) :: loss(*)
                        integer(c_int) :: rc
                end function
                function ising_mlp_forward(ctx, x, n, f) bind(C, name="ising_mlp_forward") result(rc)
                        import :: c_ptr, c_int, c_float
                        type(c_ptr), value :: ctx
                        real(c_float), intent(in) :: x(*)
                        integer(c_int), value :: n
                        real(c_float), intent(out) :: f(*)
                        integer(c_int) :: rc
                end function
                function ising_mlp_grad(ctx, records, n, step, loss, grad) bind(C, name="ising_mlp_grad") result(rc)
                        import :: c_ptr, c_int, c_int32_t, c_float, c_double
                        type(c_ptr), value :: ctx
                        real(c_float), intent(in) :: records(*)
                        integer(c_int), value :: n
                        integer(c_int32_t), value :: step
                        real(c_double), intent(out) :: loss
                        real(c_float), intent(out) :: grad(*)
                        integer(c_int) :: rc
                end function
                function ising_mfma_probe(ctx, a, b, d) bind(C, name="ising_mfma_probe") result(rc)
                        import :: c_ptr, c_int, c_float
                        type(c_ptr), value :: ctx
                        real(c_float), intent(in) :: a(*), b(*)
                        real(c_float), intent(out) :: d(*)
                        integer(c_int) :: rc
                end function
                function ising_autotune_schedule(ctx, sched, cand, n_cand, n_sweeps, max_overhead, ms) &
                                bind(C, name="ising_autotune_schedule") result(rc)
                        import :: c_ptr, c_int, c_int32_t, c_double, ising_schedule
                        type(c_ptr), value :: ctx
                        type(ising_schedule), intent(inout) :: sched
                        integer(c_int32_t), intent(in) :: cand(*)
                        integer(c_int), value :: n_cand, n_sweeps
                        real(c_double), value :: max_overhead
                        real(c_double), intent(out) :: ms(*)
                        integer(c_int) :: rc
                end function
                function ising_time_pt_sweeps(ctx, sched, n_sweeps, ms) bind(C, name="ising_time_pt_sweeps") result(rc)
                        import :: c_ptr, c_int, c_double, ising_schedule
                        type(c_ptr), value :: ctx
                        type(ising_schedule), intent(in) :: sched
                        integer(c_int), value :: n_sweeps
                        real(c_double), intent(out) :: ms
                        integer(c_int) :: rc
                end function
                function ising_time_demag(ctx, n_fields, ms) bind(C, name="ising_time_demag") result(rc)
                        import :: c_ptr, c_int, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: n_fields
                        real(c_double), intent(out) :: ms
                        integer(c_int) :: rc
                end function
                function ising_demag_field(ctx, replica, h, n) bind(C, name="ising_demag_field") result(rc)
                        import :: c_ptr, c_int, c_int64_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), intent(out) :: h(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_get_demag_field(ctx, replica, h, n) bind(C, name="ising_get_demag_field") result(rc)
                        import :: c_ptr, c_int, c_int64_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), intent(out) :: h(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_set_demag_field(ctx, replica, h, n) bind(C, name="ising_set_demag_field") result(rc)
                        import :: c_ptr, c_int, c_int64_t, c_double
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: replica
                        real(c_double), intent(in) :: h(*)
                        integer(c_int64_t), value :: n
                        integer(c_int) :: rc
                end function
                function ising_set_demag_interval(ctx, sweeps) bind(C, name="ising_set_demag_interval") result(rc)
                        import :: c_ptr, c_int
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: sweeps
                        integer(c_int) :: rc
                end function
                function c_strlen(s) bind(C, name="strlen") result(n)
                        import :: c_ptr, c_size_t
                        type(c_ptr), value :: s
                        integer(c_size_t) :: n
                end function
        end interface

contains

        ! Text of ising_last_error(ctx) as a Fortran string.
        function ising_error_text(ctx) result(msg)
                type(c_ptr), intent(in) :: ctx
                character(len=:), allocatable :: msg
                type(c_ptr) :: p
                character(kind=c_char), pointer :: chars(:)
                integer :: n, i
                p = ising_last_error(ctx)
                if (.not. c_associated(p)) then
                        msg = ""
                        return
                end if
                n = int(c_strlen(p))
                call c_f_pointer(p, chars, [n])
                allocate(character(len=n) :: msg)
                do i = 1, n
                        msg(i:i) = chars(i)
                end do
        end function ising_error_text

        ! Stop the program with the library's message when rc is an error,
        ! as the reference stops on its own failures (error stop).
        subroutine ising_check(rc, ctx, where)
                use iso_fortran_env, only: error_unit
                integer(c_int), intent(in) :: rc
                type(c_ptr), intent(in) :: ctx
                character(len=*), intent(in) :: where
                if (rc /= ISING_OK) then
                        write(error_unit, '(A,A,I0,A,A)') where, ": error ", rc, ": ", ising_error_text(ctx)
                        error stop "ising_gpu call failed"
                end if
        end subroutine ising_check
end module ising_gpu_c
