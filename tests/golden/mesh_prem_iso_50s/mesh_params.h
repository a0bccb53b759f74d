! Proc   1: Header for mesh information to run static solver
! created by the mesher on 12/03/2026, at 03h 19min
 
!:::::::::::::::::::: Input parameters :::::::::::::::::::::::::::
!   Background model     :            prem_iso
!   Dominant period [s]  :   50.0000
!   Elements/wavelength  :    1.5000
!   Courant number       :    0.6000
!   Coarsening levels    :         3
!:::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::
 
 integer, parameter ::         npol =         4  !            polynomial order
 integer, parameter ::        nelem =      6050  !                   proc. els
 integer, parameter ::       npoint =    151250  !               proc. all pts
 integer, parameter ::    nel_solid =      5258  !             proc. solid els
 integer, parameter ::    nel_fluid =       792  !             proc. fluid els
 integer, parameter :: npoint_solid =    131450  !             proc. solid pts
 integer, parameter :: npoint_fluid =     19800  !             proc. fluid pts
 integer, parameter ::  nglob_fluid =     12989  !            proc. flocal pts
 integer, parameter ::     nel_bdry =       132  ! proc. solid-fluid bndry els
 integer, parameter ::        ndisc =        12  !   # disconts in bkgrd model
 integer, parameter ::   nproc_mesh =         1  !        number of processors
 integer, parameter :: lfbkgrdmodel =         8  !   length of bkgrdmodel name
 
!:::::::::::::::::::: Output parameters ::::::::::::::::::::::::::
!   Time step [s]        :    0.1432
!   Min(h/vp),dt/courant :    0.8050    0.9548
!   max(h/vs),T0/wvlngth :   32.9845   33.3333
!   Inner core r_min [km]:  856.4651
!   Max(h) r/ns(icb) [km]:   87.2149
!   Max(h) precalc.  [km]:   89.2151
!:::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::
 
