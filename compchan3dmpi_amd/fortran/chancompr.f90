!=======================================================================
! chanCompr -- drop-in driver for the compressible channel DNS whose time
! step runs on one MI355X through the C ABI of include/cc_abi.h.
!
! File contracts are the reference's (theparvezahmad/compChan3Dmpi, v1.9):
!   in : ../param.dat (comdata.f90:31-68), the solution file named there
!        (resOp=1, main_v1.9.f90:222-251) or, resOp=2, the per-rank files
!        ../output/3Dnnn.dat of the topology in param.dat (:253-270), checked
!        against domInfo.txt where it exists
!   out: ../output/3Dnnn.dat / 3Dnnn_.dat, one per rank of that topology,
!        alternating every 1000 steps, and domInfo.txt (main_v1.9.f90:749-794);
!        Mesh.dat when debugOp = 1 (comdata.f90:112-130); KE.dat, rhobulk.dat,
!        Ubulk.dat, massflow.dat every 10 steps (:799-906); tauWall.dat,
!        moment1st/2nd/Joint.dat every istat steps (:910-1037); the 10-step
!        timing line on stdout (:1042-1047); on rho<0 or T<0 the offending
!        nodes on stdout, ../debug/3D001.dat and stop (:709-743).
! One GPU runs the whole domain: the MPI topology of param.dat only sets the
! layout of the per-rank files, exactly as the reference's ranks would write
! and read them (block bounds of main_v1.9.f90:110-145, ranks numbered
! row-major over the topology), so that either program restarts from the
! other's files.  The grid is read from the ZONE line(s) of the solution
! file(s) instead of the compile-time nP.  Environment (no counterpart in the reference, where
! these are compile-time): CC_WALL (a study's wall condition), CC_FASTIO
! (default 1: solution files through the native parallel reader/writer,
! byte-identical to the Fortran runtime's; 0: the Fortran runtime),
! CC_MAXSTEP (default 50000000), CC_PRECISION
! (fp64 | fp64-strict | mixed, default fp64), CC_NWRITE (default 1000).
!=======================================================================
program chanCompr
use iso_c_binding
use cc_iface
implicit none

integer :: resOp, topo(3), istat, debugOp
double precision :: Re, Mac, dt, PexVal, PeyVal, PezVal
character(len=10) :: descp_
character(len=100) :: inputPath_
character(len=:), allocatable :: descp, inputPath

integer :: n1, n2, n3, i, j, k, st, nstep, nsub, nnext, maxstep, nwrite, rc, prec
integer :: nproc, id, ib(3), ie(3), nn(3), cc(3), d
integer, allocatable :: bsz(:,:), bb(:,:), ee(:,:)
logical :: there
integer(c_int) :: ns_c, ijk(3)
double precision :: time, t_c, tStart, tEnd, tau(2), dg(16)
double precision, allocatable :: var(:,:,:,:), x1(:), x2(:), x3(:)
double precision, allocatable :: stat1(:,:), stat2(:,:), corr(:,:)
character(len=30) :: filename
character(len=64) :: env
character(len=200) :: zone
type(cc_params) :: prm
type(cc_wall) :: wspec
logical :: fastio
integer(c_int) :: wall
type(c_ptr) :: h
integer(8) :: clk0, clk1, crate
logical :: isBlown, isBad

!------------------------------- param.dat -------------------------------
open(unit=10, file="../param.dat", status="old")
read(10,*)
read(10,*)
read(10,'(48X,A)') descp_
read(10,'(48X,i1)') resOp
read(10,*)
read(10,'(48X,i2)') topo(1)
read(10,'(48X,i2)') topo(2)
read(10,'(48X,i2)') topo(3)
read(10,*)
read(10,*)
read(10,'(48X,F4.0)') Re
read(10,'(48X,f3.1)') Mac
read(10,'(48X,f6.4)') dt
read(10,'(48X,i2)') istat
read(10,*)
read(10,'(48X,f3.1)') PexVal
read(10,'(48X,f3.1)') PeyVal
read(10,'(48X,f3.1)') PezVal
read(10,*)
read(10,'(48X,i1)') debugOp
read(10,'(48X,A)') inputPath_
close(10)
allocate(character(len=len(trim(descp_))) :: descp)
allocate(character(len=len(trim(inputPath_))) :: inputPath)
descp = trim(descp_)
inputPath = trim(inputPath_)

maxstep = 50000000
call get_environment_variable("CC_MAXSTEP", env, status=st)
if (st == 0) read(env,*) maxstep
nwrite = 1000
call get_environment_variable("CC_NWRITE", env, status=st)
if (st == 0) read(env,*) nwrite
fastio = .true.
call get_environment_variable("CC_FASTIO", env, status=st)
if (st == 0) fastio = trim(env) /= "0"
prec = CC_FP64
call get_environment_variable("CC_PRECISION", env, status=st)
if (st == 0) then
	select case (trim(env))
	case ("fp64-strict")
		prec = CC_FP64_STRICT
	case ("mixed")
		prec = CC_MIXED
	case ("fp64")
		prec = CC_FP64
	case default
		write(6,'(3A)') "Error in CC_PRECISION: unknown mode '", trim(env), &
			"' (fp64, fp64-strict or mixed)"
		stop 1
	end select
endif

if (any(topo < 1)) then
	write(6,'(A,3(I3,1X))') "Error in param.dat: process topology must be positive: ",topo
	stop 1
endif
nproc = topo(1)*topo(2)*topo(3)

write(6,'(A)')"==================================================================================="
write(6,'(2A)')"Compressible Channel Flow Solver started at :",stamp()
write(6,'(A)')"==================================================================================="
write(6,'(A)')
write(6,'(A)')"==================================================================================="
write(6,'(A)')"Parameters read from param.dat"
write(6,'(A)')"-----------------------------------------------------------------------------------"
write(6,'(2A)')"Case Description        : ",descp
write(6,'(2A)')"Path to input file      : ",inputPath
write(6,'(A,3(I3,1X))')"Process Topology        : ",topo(1),topo(2),topo(3)
write(6,'(A,F10.4)'   )"Reynolds Number         : ",Re
write(6,'(A,F10.2)'   )"Mach Number             : ",Mac
write(6,'(A,F10.8)'   )"Timestep                : ",dt
write(6,'(A,I10)'     )"Freq of stat calculation: ",istat
write(6,'(A,F10.4)'   )"Peclet in x-direction   : ",PexVal
write(6,'(A,F10.4)'   )"Peclet in y-direction   : ",PeyVal
write(6,'(A,F10.4)'   )"Peclet in z-direction   : ",PezVal
write(6,'(A,I10)'     )"Restart Option  [1/2]   : ",resOp
write(6,'(A,I10)'     )"Debug Option    [0/1]   : ",debugOp
write(6,'(A)')"==================================================================================="

!----------------------------- solution file -----------------------------
if (resOp == 2) then
	! the per-rank files of the topology in param.dat (main_v1.9.f90:253-270):
	! the grid is the sum of the blocks along each axis, every block must be the
	! one the reference's decomposition gives its rank, and domInfo.txt, when
	! present, must list the same layout
	allocate(bsz(3,0:nproc-1), bb(3,0:nproc-1), ee(3,0:nproc-1))
	do id = 0, nproc-1
		write(filename,'(a,i3.3,a)') "../output/3D", id+1, ".dat"
		inquire(file=trim(filename), exist=there)
		if (.not. there) call topology_error(id, "missing "//trim(filename))
		open(unit=10, file=trim(filename), status="old")
		if (id == 0) then
			read(10,'(8X,F20.10,I9,1X)') time, nstep
		else
			read(10,*)
		endif
		read(10,*)
		read(10,'(A)') zone
		close(10)
		read(zone,'(7X,I3,3X,I3,3X,I3)', iostat=st) bsz(3,id), bsz(2,id), bsz(1,id)
		if (st /= 0) call topology_error(id, "no ZONE line in "//trim(filename))
	enddo
	nn = 0
	do d = 1, 3
		do i = 0, topo(d)-1
			cc = 0
			cc(d) = i
			nn(d) = nn(d) + bsz(d, rank_of(cc))
		enddo
	enddo
	n1 = nn(1); n2 = nn(2); n3 = nn(3)
	do id = 0, nproc-1
		call block_of(nn, id, bb(:,id), ee(:,id))
		if (any(ee(:,id) - bb(:,id) + 1 /= bsz(:,id))) &
			call topology_error(id, "its block does not fit the topology in param.dat")
	enddo
	inquire(file="../output/domInfo.txt", exist=there)
	if (there) then
		open(unit=10, file="../output/domInfo.txt", status="old")
		read(10,105,iostat=st) ib
		if (st == 0) read(10,106,iostat=st) i
		if (st /= 0 .or. any(ib /= nn) .or. i /= nproc) &
			call topology_error(0, "domInfo.txt lists another grid or rank count")
		do id = 0, nproc-1
			read(10,107,iostat=st) ib(1), ie(1), ib(2), ie(2), ib(3), ie(3)
			if (st /= 0 .or. any(ib /= bb(:,id)) .or. any(ie /= ee(:,id))) &
				call topology_error(id, "domInfo.txt lists another block")
		enddo
		close(10)
	endif
	allocate(var(n1,n2,n3,7), x1(n1), x2(n2), x3(n3))
	do id = 0, nproc-1
		write(filename,'(a,i3.3,a)') "../output/3D", id+1, ".dat"
		call read_block(trim(filename), id, bb(:,id), ee(:,id))
	enddo
else
	open(unit=10, file=inputPath, status="old")
	read(10,'(8X,F20.10,I9,1X)') time, nstep
	read(10,*)
	read(10,'(A)') zone
	read(zone,'(7X,I3,3X,I3,3X,I3)', iostat=st) n3, n2, n1
	if (st /= 0) then
		write(6,'(A)') "Error in reading the ZONE line of the input file"
		stop
	endif
	close(10)
	allocate(var(n1,n2,n3,7), x1(n1), x2(n2), x3(n3))
	call read_block(inputPath, -1, [1,1,1], [n1,n2,n3])
	allocate(bb(3,0:nproc-1), ee(3,0:nproc-1))
	do id = 0, nproc-1
		call block_of([n1,n2,n3], id, bb(:,id), ee(:,id))
	enddo
endif
allocate(stat1(n3,7), stat2(n3,7), corr(n3,9))
write(6,'(A)') "Input file successfully read"
if (debugOp == 1) call write_mesh()

!--------------------------- device instance -----------------------------
prm%n1 = n1; prm%n2 = n2; prm%n3 = n3
prm%istat = istat; prm%precision = prec; prm%reserved = 0
prm%Re = Re; prm%Mac = Mac; prm%dt = dt
prm%PexVal = PexVal; prm%PeyVal = PeyVal; prm%PezVal = PezVal
prm%Ri_b = 0.0d0; prm%slip_len = 0.0d0
rc = cc_create(prm, h)
if (rc /= 0) call fatal("cc_create")
! CC_WALL=<study> applies the wall condition of one of the reference's
! parameter studies (v1.9_STWNV s1..s6, v1.9_RD rd1..rd3, v1.9_groov g1,
! v1.9_groov_strat gst1..gst5, v1.9_strat st1..st4) at the file's x1, x2
call get_environment_variable("CC_WALL", env, status=st)
if (st == 0 .and. len_trim(env) > 0) then
	rc = cc_wall_preset(trim(env)//c_null_char, wall, wspec)
	if (rc /= 0) call fatal("cc_wall_preset")
	rc = cc_set_wall(h, wall, wspec, x1, x2)
	if (rc /= 0) call fatal("cc_set_wall")
	write(6,'(2A)') "Wall condition of study ", trim(env)
endif
rc = cc_upload(h, var, int(nstep, c_int), time)
if (rc /= 0) call fatal("cc_upload")

!------------------------------- time loop -------------------------------
call system_clock(clk0, crate)
do while (nstep < maxstep)
	! advance to the next step at which the reference does any output
	nnext = min(maxstep, (nstep/10 + 1)*10, (nstep/istat + 1)*istat, (nstep/nwrite + 1)*nwrite)
	nsub = nnext - nstep
	if (mod(nstep,10) == 0) call system_clock(clk0)
	rc = cc_step(h, int(nsub, c_int))
	if (rc < 0) call fatal("cc_step")
	isBlown = rc > 0
	if (isBlown) nsub = rc - nstep
	do i = 1, nsub
		nstep = nstep + 1
		time = time + dt
	enddo

	!----------------------- check for unphysical values ----------------------
	if (isBlown) then
		rc = cc_download(h, var, ns_c, t_c)
		if (rc /= 0) call fatal("cc_download")
		! each rank's block: its offending nodes on stdout and, when it has
		! any, its debug file (main_v1.9.f90:728-766)
		do id = 0, nproc-1
			isBad = .false.
			do k=bb(3,id),ee(3,id)
				do j=bb(2,id),ee(2,id)
					do i=bb(1,id),ee(1,id)
						if (var(i,j,k,1) < 0.0d0 .or. var(i,j,k,6) < 0.0d0) then
							write(6,102) nstep,id,i,j,k,var(i,j,k,1),var(i,j,k,6)
							isBad = .true.
						endif
					enddo
				enddo
			enddo
			if (.not. isBad) cycle
			write(filename,'(a,i3.3,a)') "../debug/3D", id+1, ".dat"
			open(unit=10, file=trim(filename))
			write(10,103) 'TITLE ="',time,nstep,'"'
			write(10,*) 'Variables ="x","y","z","Rho","U","V","W","E","T","P"'
			write(10,104) 'ZONE k=',ee(3,id)-bb(3,id)+1,',j=',ee(2,id)-bb(2,id)+1,',i=',ee(1,id)-bb(1,id)+1, &
				',DATAPACKING="POINT"'
			do k=bb(3,id),ee(3,id)
				write(10,*)
				do j=bb(2,id),ee(2,id)
					write(10,*)
					do i=bb(1,id),ee(1,id)
						write(10,116) x1(i),x2(j),x3(k),var(i,j,k,1),var(i,j,k,2),var(i,j,k,3),var(i,j,k,4), &
							var(i,j,k,5),var(i,j,k,6),var(i,j,k,7)
					enddo
				enddo
			enddo
			close(10)
		enddo
		call cc_destroy(h)
		stop
	endif

	!--------------------------- solution file ---------------------------
	if (mod(nstep,nwrite) == 0) then
		rc = cc_download(h, var, ns_c, t_c)
		if (rc /= 0) call fatal("cc_download")
		! one file per rank of the topology, as the ranks write them
		do id = 0, nproc-1
			if (mod(nstep/nwrite,2) == 0) then
				write(filename,'(a,i3.3,a)') "../output/3D", id+1, ".dat"
			else
				write(filename,'(a,i3.3,a)') "../output/3D", id+1, "_.dat"
			endif
			call write_block(trim(filename), bb(:,id), ee(:,id))
		enddo
		open(unit=10, file="../output/domInfo.txt")
		write(10,105) n1, n2, n3
		write(10,106) nproc
		do id = 0, nproc-1
			write(10,107) bb(1,id), ee(1,id), bb(2,id), ee(2,id), bb(3,id), ee(3,id)
		enddo
		close(10)
	endif

	!------------------------- bulk diagnostics --------------------------
	if (mod(nstep,10) == 0) then
		rc = cc_diag(h, x3, x2(n2), dg)
		if (rc /= 0) call fatal("cc_diag")
		open(10, file="../output/KE.dat", access="append")
		write(10,108) time, dg(1)
		close(10)
		open(10, file="../output/rhobulk.dat", access="append")
		write(10,109) time, dg(2:6)
		close(10)
		open(10, file="../output/Ubulk.dat", access="append")
		write(10,109) time, dg(7:11)
		close(10)
		open(10, file="../output/massflow.dat", access="append")
		write(10,109) time, dg(12:16)
		close(10)
	endif

	!---------------------- wall shear and moments -----------------------
	if (mod(nstep,istat) == 0) then
		rc = cc_stats(h, tau, stat1, stat2, corr)
		if (rc /= 0) call fatal("cc_stats")
		open(10, file="../output/tauWall.dat", access="append")
		write(10,110) time, tau(1), tau(2)
		close(10)
		open(unit=10, file="../output/moment1st.dat", access="append")
		write(10,111,advance="no") time
		do k=1,n3
			write(10,112,advance="no") stat1(k,1),stat1(k,2),stat1(k,3),stat1(k,4),stat1(k,5),stat1(k,6),stat1(k,7)
		enddo
		write(10,*)
		close(10)
		open(unit=10, file="../output/moment2nd.dat", access="append")
		write(10,111,advance="no") time
		do k=1,n3
			write(10,112,advance="no") stat2(k,1),stat2(k,2),stat2(k,3),stat2(k,4),stat2(k,5),stat2(k,6),stat2(k,7)
		enddo
		write(10,*)
		close(10)
		open(unit=10, file="../output/momentJoint.dat", access="append")
		write(10,111,advance="no") time
		do k=1,n3
			write(10,113,advance="no") corr(k,1),corr(k,2),corr(k,3),corr(k,4),corr(k,5),corr(k,6), &
				corr(k,7),corr(k,8),corr(k,9)
		enddo
		write(10,*)
		close(10)
	endif

	if (mod(nstep,10) == 0) then
		call system_clock(clk1)
		tStart = 0.0d0
		tEnd = dble(clk1 - clk0)/dble(crate)
		write(6,114) nstep,time," | CPU Time:",tEnd-tStart,"sec for 10 iters | ", stamp()
		write(6,'(A)')"-----------------------------------------------------------------------------------"
	endif
enddo

call cc_destroy(h)
write(*,*) "Compressible Channel Flow Solver exited at :",stamp()

101 format(3(1X,F10.6),7(1X,E22.15))
102 format(I10,4(I4),2(F8.4))
103 format(A,F20.10,I9,A)
104 format(A,I3,A,I3,A,I3,A)
105 format(3(i4))
106 format(i3)
107 format(6(i4,1x))
108 format(2(1X,F15.7))
109 format(6(1X,F15.7))
110 format(3(1X,F15.7))
111 format(1(1X,F15.7))
112 format(7(1X,F15.7))
113 format(9(1X,F15.7))
114 format(I10,F10.5,A,F6.3,2(1X,A))
116 format(3(1X,F7.4),7(1X,E12.5))

contains

! rank of the topology coordinates c (0-based): row-major, the last axis
! fastest, as MPI numbers a Cartesian communicator created without reordering
! (main_v1.9.f90:90)
integer function rank_of(c)
	integer, intent(in) :: c(3)
	rank_of = (c(1)*topo(2) + c(2))*topo(3) + c(3)
end function rank_of

! bounds of rank id's block of an n(1) x n(2) x n(3) grid (main_v1.9.f90:110-145):
! along each axis the first topo - mod(n, topo) ranks take n/topo nodes and
! the rest one more
subroutine block_of(n, id, lo, hi)
	integer, intent(in) :: n(3), id
	integer, intent(out) :: lo(3), hi(3)
	integer :: c(3), dd, r, sz, first
	c(3) = mod(id, topo(3))
	c(2) = mod(id/topo(3), topo(2))
	c(1) = id/(topo(3)*topo(2))
	do dd = 1, 3
		sz = n(dd)/topo(dd)
		first = 1
		do r = 0, topo(dd)-1
			if (r == topo(dd) - mod(n(dd), topo(dd))) sz = sz + 1
			if (r == c(dd)) then
				lo(dd) = first
				hi(dd) = first + sz - 1
			endif
			first = first + sz
		enddo
	enddo
end subroutine block_of

! the reference's message for a restart whose files do not fit the topology
! (main_v1.9.f90:264-266)
subroutine topology_error(id, what)
	integer, intent(in) :: id
	character(len=*), intent(in) :: what
	write(6,'(A,1X,I4,2A)') "Error in reading component input file at", id, ": ", what
	write(6,'(A)') "Change in MPI process topology since last run could possibly be a reason"
	stop 1
end subroutine topology_error

! one solution file into var(lo:hi); id >= 0 for a rank's file (resOp=2),
! -1 for the combined file (resOp=1)
subroutine read_block(path, id, lo, hi)
	character(len=*), intent(in) :: path
	integer, intent(in) :: id, lo(3), hi(3)
	integer :: s1, s2, s3, ii, jj, kk, ios
	double precision, allocatable :: vb(:,:,:,:), y1(:), y2(:), y3(:)
	s1 = hi(1)-lo(1)+1; s2 = hi(2)-lo(2)+1; s3 = hi(3)-lo(3)+1
	if (fastio) then
		allocate(vb(s1,s2,s3,7), y1(s1), y2(s2), y3(s3))
		ios = cc_read_solution(path//c_null_char, s1, s2, s3, y1, y2, y3, vb, t_c, ns_c, 0)
		if (ios /= 0) then
			if (id >= 0) call topology_error(id, cc_ioerrmsg())
			write(6,'(2A)') "Error in reading the input file: ", cc_ioerrmsg()
			stop
		endif
		var(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3),:) = vb
		x1(lo(1):hi(1)) = y1
		x2(lo(2):hi(2)) = y2
		x3(lo(3):hi(3)) = y3
		return
	endif
	open(unit=10, file=path, status="old")
	read(10,*)
	read(10,*)
	read(10,*)
	do kk=lo(3),hi(3)
		read(10,*)
		do jj=lo(2),hi(2)
			read(10,*)
			do ii=lo(1),hi(1)
				read(10,'(3(1X,F10.6),7(1X,E22.15))',iostat=ios) x1(ii),x2(jj),x3(kk),var(ii,jj,kk,1), &
					var(ii,jj,kk,2),var(ii,jj,kk,3),var(ii,jj,kk,4),var(ii,jj,kk,5),var(ii,jj,kk,6),var(ii,jj,kk,7)
				if (ios /= 0) then
					if (id >= 0) then
						write(6,'(A,5(1X,I4))') "Error in reading component input file at",id,ii,jj,kk,ios
						write(6,'(A)') "Change in MPI process topology since last run could possibly be a reason"
					else
						write(6,'(A,4(1X,I4))') "Error in reading combined input file at",ii,jj,kk,ios
					endif
					stop
				endif
			enddo
		enddo
	enddo
	close(10)
end subroutine read_block

! var(lo:hi) as one solution file (main_v1.9.f90:759-773)
subroutine write_block(path, lo, hi)
	character(len=*), intent(in) :: path
	integer, intent(in) :: lo(3), hi(3)
	integer :: s1, s2, s3, ii, jj, kk, ios
	double precision, allocatable :: vb(:,:,:,:), y1(:), y2(:), y3(:)
	s1 = hi(1)-lo(1)+1; s2 = hi(2)-lo(2)+1; s3 = hi(3)-lo(3)+1
	if (fastio) then
		allocate(vb(s1,s2,s3,7), y1(s1), y2(s2), y3(s3))
		vb = var(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3),:)
		y1 = x1(lo(1):hi(1)); y2 = x2(lo(2):hi(2)); y3 = x3(lo(3):hi(3))
		ios = cc_write_solution(path//c_null_char, time, int(nstep, c_int), s1, s2, s3, &
			y1, y2, y3, vb, c_null_ptr, 0)
		if (ios /= 0) then
			write(6,'(2A)') "Error in writing the solution file: ", cc_ioerrmsg()
			stop
		endif
		return
	endif
	open(unit=10, file=path)
	write(10,'(A,F20.10,I9,A)') 'TITLE ="',time,nstep,'"'
	write(10,*) 'Variables ="x","y","z","Rho","U","V","W","E","T","P"'
	write(10,'(A,I3,A,I3,A,I3,A)') 'ZONE k=',s3,',j=',s2,',i=',s1,',DATAPACKING="POINT"'
	do kk=lo(3),hi(3)
		write(10,*)
		do jj=lo(2),hi(2)
			write(10,*)
			do ii=lo(1),hi(1)
				write(10,'(3(1X,F10.6),7(1X,E22.15))') x1(ii),x2(jj),x3(kk),var(ii,jj,kk,1),var(ii,jj,kk,2),var(ii,jj,kk,3), &
					var(ii,jj,kk,4),var(ii,jj,kk,5),var(ii,jj,kk,6),var(ii,jj,kk,7)
			enddo
		enddo
	enddo
	close(10)
end subroutine write_block

! Mesh.dat (comdata.f90:112-130): the analytic mesh of setupMesh
! (comdata.f90:70-110), written before the input file overwrites x1..x3
subroutine write_mesh()
	double precision, allocatable :: m1(:), m2(:), m3(:), xi(:)
	double precision :: pi_, dx_, dy_, dxi_
	integer :: ii, jj, kk
	allocate(m1(n1), m2(n2), m3(n3), xi(n3))
	pi_ = 4.d0*datan(1.0d0)
	m1(1) = 0.0d0; m2(1) = 0.0d0; xi(1) = 0.0d0; m3(1) = 0.0d0
	m1(n1) = 4.0d0*pi_
	m2(n2) = 4.0d0*pi_/3.0d0
	m3(n3) = 2.0d0
	dx_ = m1(n1)/dble(n1-1)
	dy_ = m2(n2)/dble(n2-1)
	dxi_ = m3(n3)/dble(n3-1)
	do ii = 1, n1-1
		m1(ii+1) = m1(ii) + dx_
	enddo
	do jj = 1, n2-1
		m2(jj+1) = m2(jj) + dy_
	enddo
	do kk = 1, n3-1
		xi(kk+1) = xi(kk) + dxi_
	enddo
	do kk = 1, n3
		m3(kk) = 1.0d0 - dcos(0.50d0*xi(kk)*pi_)
	enddo
	open(unit=10, file='Mesh.dat')
	write(10,*) 'TITLE ="Mesh"'
	write(10,*) 'Variables ="x","y","z"'
	write(10,'(3(A,I3),A)') 'ZONE k=',n3,',j=',n2,',i=',n1,',DATAPACKING="POINT"'
	do kk = 1, n3
		write(10,*)
		do jj = 1, n2
			write(10,*)
			do ii = 1, n1
				write(10,'(3(F6.4))') m1(ii), m2(jj), m3(kk)
			enddo
		enddo
	enddo
	close(10)
end subroutine write_mesh

subroutine fatal(where)
	character(len=*), intent(in) :: where
	write(6,'(4A)') "Error in ", where, ": ", cc_errmsg()
	stop 1
end subroutine fatal

! wall-clock stamp "d Mon yyyy  hh:mm:ss.mmm AM|PM" (the reference's layout)
function stamp() result(s)
	character(len=30) :: s
	character(len=3), parameter :: mon(12) = ['Jan','Feb','Mar','Apr','May','Jun', &
		'Jul','Aug','Sep','Oct','Nov','Dec']
	integer :: v(8), hh
	character(len=3) :: ap
	call date_and_time(values=v)
	hh = v(5)
	ap = 'AM'
	if (hh >= 12) ap = 'PM'
	if (hh > 12) hh = hh - 12
	write(s,'(i2,1x,a,1x,i4,2x,i2,a1,i2.2,a1,i2.2,a1,i3.3,1x,a)') v(3), mon(v(2)), v(1), hh, ':', &
		v(6), ':', v(7), '.', v(8), trim(ap)
end function stamp

end program chanCompr
