"""ctypes declarations of the library's recon / learned-prox / batch / host-level entry points."""
import ctypes as C

_vp = C.c_void_p
_fp = C.c_void_p      # device pointers are passed as integers
_dp = C.POINTER(C.c_double)
_hf = C.POINTER(C.c_float)    # host float arrays
_hb = C.POINTER(C.c_ubyte)    # host byte arrays


def declare(L):
    i, f, d, ll = C.c_int, C.c_float, C.c_double, C.c_longlong
    sig = {
        "ctlab_recon_create": ([_vp, i], _vp),
        "ctlab_recon_destroy": ([_vp], None),
        "ctlab_recon_set_data": ([_vp, _fp, _vp], i),
        "ctlab_recon_set_tiles": ([_vp, i, i], i),
        "ctlab_os_sart": ([_vp, _fp, i, f, i, i, _vp], i),
        "ctlab_fista_tv": ([_vp, _fp, i, f, i, f, i, _dp, _vp], i),
        "ctlab_cg": ([_vp, _fp, i, i, _dp, _vp], i),
        "ctlab_cg_to_fidelity": ([_vp, _fp, i, d, _dp, C.POINTER(i), _vp], i),
        "ctlab_recon_data_fidelity": ([_vp, _fp, _vp], d),
        "ctlab_recon_gradient": ([_vp, _fp, _fp, _vp], d),
        "ctlab_recon_lipschitz": ([_vp, i, _vp], d),
        "ctlab_sart_prj_weight": ([_fp, _fp, _fp, ll, _vp], i),
        "ctlab_sart_bak_weight": ([_fp, _fp, _fp, ll, _vp], i),
        "ctlab_ossart_update": ([_fp, _fp, _fp, _fp, f, ll, _vp], i),
        "ctlab_fista_momentum": ([_fp, _fp, f, f, f, f, ll, _vp], i),
        "ctlab_soft_threshold": ([_fp, f, ll, _vp], i),
        "ctlab_axpby": ([f, _fp, f, _fp, ll, _vp], i),
        "ctlab_cg_beta": ([_fp, _fp, _fp, i, ll, _dp, _vp], i),
        "ctlab_cg_beta_static": ([_fp, _fp, _fp, i, ll, _vp, _vp], i),
        "ctlab_tv_prox": ([_fp, _fp, f, i, i, i, i, _fp, _vp], i),
        "ctlab_batch_create": ([C.POINTER(_vp), i], _vp),
        "ctlab_batch_destroy": ([_vp], None),
        "ctlab_batch_size": ([_vp], i),
        "ctlab_batch_set_tiles": ([_vp, i, i], i),
        "ctlab_batch_get_tiles": ([_vp, C.POINTER(i), C.POINTER(i)], i),
        "ctlab_batch_os_sart": ([_vp, C.POINTER(_vp), i, f, i, i, _vp], i),
        "ctlab_batch_autotune": ([_vp, C.POINTER(_vp), f, C.POINTER(f)], i),
        "ctlab_cnn_num_params": ([], i),
        "ctlab_cnn_create": ([i, i, i, i, i], _vp),
        "ctlab_cnn_destroy": ([_vp], None),
        "ctlab_cnn_set_params": ([_vp, C.POINTER(C.c_float)], i),
        "ctlab_cnn_get_params": ([_vp, C.POINTER(C.c_float)], i),
        "ctlab_cnn_forward": ([_vp, _fp, _fp, _vp], i),
        "ctlab_cnn_loss_grad": ([_vp, _fp, _fp, _fp, C.POINTER(C.c_float), _vp], d),
        "ctlab_cnn_adam_step": ([_vp, f, f, f, f, _vp], i),
        "ctlab_getVolZIdxPair": ([_hf, i, i, i, f, f, i, f, f, i, C.POINTER(i), C.POINTER(i)], i),
        "ctlab_getPrjIdxPair": ([_hf, i, i, i, f, f, i, f, f, i, C.POINTER(i), C.POINTER(i)], i),
        "ctlab_read_raw": ([C.c_char_p, _vp, ll, i], i),
        "ctlab_write_raw": ([C.c_char_p, _vp, ll, i], i),
        "ctlab_read_raw_device": ([C.c_char_p, _fp, ll, i], i),
        "ctlab_write_raw_device": ([C.c_char_p, _fp, ll, i], i),
        "ctlab_evalProjBack": ([_vp, i, i, i, _fp, _vp], i),
        "ctlab_arc_detector": ([f, f, i, i, f, f, f, f, _hf, _hf, _hf], i),
        "ctlab_OS_SART": ([_hf, _hf, _hf, _hf, _hf, _hb, C.c_char_p, i, i, i, f, f, i, i, i, f, f, f, i, i, i, f, f,
                           f, f, f, f, i, i], i),
        "ctlab_CG": ([_hf, _hf, _hf, _hf, _hf, _hb, C.c_char_p, i, i, i, f, f, i, i, i, f, f, f, i, i, i, f, f,
                      f, f, f, f, i, i, i], i),
    }
    for name, (args, res) in sig.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
