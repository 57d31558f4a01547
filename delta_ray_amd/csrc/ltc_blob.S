/* The fitted LTC matrices (16384 x 9 float32, delta_ray_amd/data/ltc_tabM.f32)
   embedded read-only in the library, so no data file travels with it. */
    .section .rodata
    .global dr_ltc_blob
    .global dr_ltc_blob_end
    .balign 64
dr_ltc_blob:
    .incbin LTC_FILE
dr_ltc_blob_end:
    .section .note.GNU-stack,"",@progbits
