//! Rust binding of the MI355X path tracer (include/hpt.h).
//!
//! `HalaRenderer` keeps the method names the reference host calls on
//! `hala_renderer::rt_renderer::HalaRenderer` (src/bin/hala-pathtracer.rs:128-285);
//! the Vulkan-only calls (`push_*_shader_with_file`, the imgui context) have no
//! counterpart, and the shader permutation becomes the `features` bits.
//! Errors map a non-zero status to `anyhow!(hpt_last_error())`.
//!
//! Written against the C header; no Rust toolchain was available where it was
//! written, so it has not been compiled.
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};

pub type hpt_status = i32;

pub const HPT_FEATURE_MEDIUM: u32 = 1;
pub const HPT_FEATURE_VOL_MIS: u32 = 2;
pub const HPT_FEATURE_TRANSPARENT: u32 = 4;

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_camera {
    pub position: [f32; 3],
    pub _pad0: f32,
    pub right: [f32; 3],
    pub _pad1: f32,
    pub up: [f32; 3],
    pub _pad2: f32,
    pub forward: [f32; 3],
    pub yfov: f32,
    pub focal_distance_or_xmag: f32,
    pub aperture_or_ymag: f32,
    pub camera_type: u32,
    pub _pad3: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_light {
    pub intensity: [f32; 3],
    pub _pad0: f32,
    pub position: [f32; 3],
    pub _pad1: f32,
    pub u: [f32; 3],
    pub _pad2: f32,
    pub v: [f32; 3],
    pub radius: f32,
    pub area: f32,
    pub light_type: i32,
    pub _pad3: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_medium {
    pub color: [f32; 3],
    pub density: f32,
    pub anisotropy: f32,
    pub medium_type: u32,
    pub _padding0: f32,
    pub _padding1: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_material {
    pub medium: hpt_medium,
    pub base_color: [f32; 3],
    pub opacity: f32,
    pub emission: [f32; 3],
    pub anisotropic: f32,
    pub metallic: f32,
    pub roughness: f32,
    pub subsurface: f32,
    pub specular_tint: f32,
    pub sheen: f32,
    pub sheen_tint: f32,
    pub clearcoat: f32,
    pub clearcoat_roughness: f32,
    pub clearcoat_tint: [f32; 3],
    pub specular_transmission: f32,
    pub ior: f32,
    pub ax: f32,
    pub ay: f32,
    pub base_color_map_index: u32,
    pub normal_map_index: u32,
    pub metallic_roughness_map_index: u32,
    pub emission_map_index: u32,
    pub material_type: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_vertex {
    pub position: [f32; 3],
    pub normal: [f32; 3],
    pub tangent: [f32; 3],
    pub tex_coord: [f32; 2],
}

#[repr(C)]
pub struct hpt_primitive_desc {
    pub transform: [f32; 16],
    pub material_index: u32,
    pub vertex_count: u32,
    pub index_count: u32,
    pub _pad: u32,
    pub vertices: *const hpt_vertex,
    pub indices: *const u32,
}

#[repr(C)]
pub struct hpt_texture_desc {
    pub width: u32,
    pub height: u32,
    pub rgba8: *const u8,
}

#[repr(C)]
pub struct hpt_scene_desc {
    pub cameras: *const hpt_camera,
    pub camera_count: u32,
    pub light_count: u32,
    pub lights: *const hpt_light,
    pub materials: *const hpt_material,
    pub material_count: u32,
    pub primitive_count: u32,
    pub primitives: *const hpt_primitive_desc,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct hpt_create_info {
    pub width: u32,
    pub height: u32,
    pub max_depth: u32,
    pub rr_depth: u32,
    pub enable_tonemap: u32,
    pub enable_aces: u32,
    pub use_simple_aces: u32,
    pub features: u32,
    pub max_samples: u64,
    pub exposure_value: f32,
    pub device: i32,
}

#[repr(C)]
pub struct hpt_renderer {
    _private: [u8; 0],
}

#[repr(C)]
pub struct hpt_group {
    _private: [u8; 0],
}

pub const HPT_IMAGE_FINAL_RGBA8: c_int = 0;
pub const HPT_IMAGE_ACCUM_RGBA32F: c_int = 1;
pub const HPT_IMAGE_ALBEDO_RGBA32F: c_int = 2;
pub const HPT_IMAGE_NORMAL_RGBA32F: c_int = 3;
pub const HPT_LAUNCH_EAGER: u32 = 0;
pub const HPT_LAUNCH_GRAPH: u32 = 1;
pub const HPT_SHADE_SORT_OFF: u32 = 0xFFFF_FFFF;
pub const HPT_TAIL_FROM_NONE: u32 = 0xFFFF_FFFF;
pub const HPT_BVH_HOST: u32 = 0;
pub const HPT_BVH_GPU: u32 = 1;

extern "C" {
    pub fn hpt_create(info: *const hpt_create_info, out: *mut *mut hpt_renderer) -> hpt_status;
    pub fn hpt_set_scene(r: *mut hpt_renderer, scene: *const hpt_scene_desc) -> hpt_status;
    pub fn hpt_set_env_map(r: *mut hpt_renderer, rgb: *const f32, width: u32, height: u32, rotation: f32) -> hpt_status;
    pub fn hpt_exr_info(data: *const u8, size: usize, width: *mut u32, height: *mut u32) -> hpt_status;
    pub fn hpt_exr_decode_rgb(data: *const u8, size: usize, rgb: *mut f32, width: u32, height: u32) -> hpt_status;
    pub fn hpt_set_ground_color(r: *mut hpt_renderer, rgba: *const f32) -> hpt_status;
    pub fn hpt_set_sky_color(r: *mut hpt_renderer, rgba: *const f32) -> hpt_status;
    pub fn hpt_set_env_intensity(r: *mut hpt_renderer, intensity: f32) -> hpt_status;
    pub fn hpt_set_exposure_value(r: *mut hpt_renderer, ev: f32) -> hpt_status;
    pub fn hpt_set_blue_noise(r: *mut hpt_renderer, rgba8: *const u8, width: u32, height: u32) -> hpt_status;
    pub fn hpt_set_textures(r: *mut hpt_renderer, textures: *const hpt_texture_desc, count: u32) -> hpt_status;
    pub fn hpt_commit(r: *mut hpt_renderer) -> hpt_status;
    pub fn hpt_render(r: *mut hpt_renderer, frames: u32, camera_index: u32) -> hpt_status;
    pub fn hpt_reset_accumulation(r: *mut hpt_renderer) -> hpt_status;
    pub fn hpt_get_frame_index(r: *mut hpt_renderer, frame_index: *mut u64) -> hpt_status;
    pub fn hpt_set_frame_index(r: *mut hpt_renderer, frame_index: u64) -> hpt_status;
    pub fn hpt_set_frames_in_flight(r: *mut hpt_renderer, frames: u32) -> hpt_status;
    pub fn hpt_get_frames_in_flight(r: *mut hpt_renderer, frames: *mut u32) -> hpt_status;
    pub fn hpt_set_launch_mode(r: *mut hpt_renderer, mode: u32) -> hpt_status;
    pub fn hpt_set_shade_sort(r: *mut hpt_renderer, first_bounce: u32) -> hpt_status;
    pub fn hpt_set_vmis_phases(r: *mut hpt_renderer, on: u32) -> hpt_status;
    pub fn hpt_get_truncated_paths(r: *mut hpt_renderer, paths: *mut u64) -> hpt_status;
    pub fn hpt_set_tail_from(r: *mut hpt_renderer, first_bounce: u32) -> hpt_status;
    pub fn hpt_group_set_tail_from(g: *mut hpt_group, first_bounce: u32) -> hpt_status;
    pub fn hpt_set_traversal_block(r: *mut hpt_renderer, threads: u32) -> hpt_status;
    pub fn hpt_group_set_traversal_block(g: *mut hpt_group, threads: u32) -> hpt_status;
    pub fn hpt_set_traversal_refill(r: *mut hpt_renderer, on: u32) -> hpt_status;
    pub fn hpt_group_set_traversal_refill(g: *mut hpt_group, on: u32) -> hpt_status;
    pub fn hpt_group_create(device: i32, out: *mut *mut hpt_group) -> hpt_status;
    pub fn hpt_group_render(g: *mut hpt_group, jobs: *const *mut hpt_renderer, frames: *const u32,
        camera_indices: *const u32, count: u32) -> hpt_status;
    pub fn hpt_group_set_launch_mode(g: *mut hpt_group, mode: u32) -> hpt_status;
    pub fn hpt_group_set_shade_sort(g: *mut hpt_group, first_bounce: u32) -> hpt_status;
    pub fn hpt_group_get_stats(g: *mut hpt_group, steps: *mut u64, graph_launches: *mut u64,
        device_bytes: *mut u64) -> hpt_status;
    pub fn hpt_group_get_truncated_paths(g: *mut hpt_group, paths: *mut u64) -> hpt_status;
    pub fn hpt_group_wait_idle(g: *mut hpt_group) -> hpt_status;
    pub fn hpt_group_destroy(g: *mut hpt_group);
    pub fn hpt_set_bvh_builder(r: *mut hpt_renderer, builder: u32) -> hpt_status;
    pub fn hpt_get_image_pointer(r: *mut hpt_renderer, which: c_int, ptr: *mut *mut c_void, bytes: *mut usize)
        -> hpt_status;
    pub fn hpt_read_image(r: *mut hpt_renderer, which: c_int, dst: *mut c_void, bytes: usize) -> hpt_status;
    pub fn hpt_wait_idle(r: *mut hpt_renderer) -> hpt_status;
    pub fn hpt_destroy(r: *mut hpt_renderer);
    pub fn hpt_last_error() -> *const c_char;
    pub fn hpt_trace_closest_device(r: *mut hpt_renderer, rays: *const c_void, count: u32, hits: *mut c_void)
        -> hpt_status;
    pub fn hpt_present_create(r: *mut hpt_renderer, slots: u32, out: *mut *mut hpt_presenter) -> hpt_status;
    pub fn hpt_present_frame(p: *mut hpt_presenter) -> hpt_status;
    pub fn hpt_present_acquire(p: *mut hpt_presenter, pixels: *mut *const c_void, frame_index: *mut u64)
        -> hpt_status;
    pub fn hpt_present_get_stats(p: *mut hpt_presenter, presented: *mut u64, dropped: *mut u64) -> hpt_status;
    pub fn hpt_present_destroy(p: *mut hpt_presenter);
    pub fn hpt_get_image_size(r: *mut hpt_renderer, width: *mut u32, height: *mut u32) -> hpt_status;
    pub fn hpt_denoiser_create(npz: *const u8, size: usize, device: i32, out: *mut *mut hpt_denoiser)
        -> hpt_status;
    pub fn hpt_denoise(d: *mut hpt_denoiser, r: *mut hpt_renderer, dst: *mut f32, bytes: usize,
        dst_on_device: c_int) -> hpt_status;
    pub fn hpt_denoiser_destroy(d: *mut hpt_denoiser);
}

/// Opaque denoiser handle (hpt_denoiser_*).
#[repr(C)]
pub struct hpt_denoiser {
    _private: [u8; 0],
}

// ------------------------------------------------------------ serving loop
// hpt_server_* of include/hpt.h: admission by device footprint, jobs on
// streams of their own (or one wavefront per scene), results by id.

/// Opaque server handle (hpt_server_*).
#[repr(C)]
pub struct hpt_server {
    _private: [u8; 0],
}

pub const HPT_SERVE_STREAMS: u32 = 0;
pub const HPT_SERVE_WAVEFRONT: u32 = 1;
pub const HPT_SERVE_DEFAULT: u32 = 0xFFFF_FFFF;

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct hpt_server_config {
    pub device: i32,
    pub mode: u32,
    pub budget_bytes: u64,
    pub max_frames: u32,
    pub keep_images: u32,
    pub step_paths: u64,
    pub launch_mode: u32,
    pub max_running: u32,
    pub dry_run_ns_per_path: u64,
    pub reuse_handles: u32,
    pub _reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct hpt_serve_scene_desc {
    pub scene: *const hpt_renderer,
    pub features: u32,
    pub step_launch_mode: u32,
    pub step_shade_sort: u32,
    pub step_tail_from: u32,
    pub step_traversal_block: u32,
    pub dry_run_fail: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct hpt_job_desc {
    pub scene: u32,
    pub width: u32,
    pub height: u32,
    pub spp: u32,
    pub max_depth: u32,
    pub rr_depth: u32,
    pub camera_index: u32,
    pub enable_tonemap: u32,
    pub enable_aces: u32,
    pub use_simple_aces: u32,
    pub exposure_value: f32,
    pub frames_in_flight: u32,
    pub launch_mode: u32,
    pub shade_sort: u32,
    pub tail_from: u32,
    pub traversal_block: u32,
    pub denoise: u32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct hpt_job_result {
    pub job_id: u64,
    pub status: hpt_status,
    pub width: u32,
    pub height: u32,
    pub frames: u32,
    pub has_image: u32,
    pub _pad: u32,
    pub footprint_bytes: u64,
    pub submitted_s: f64,
    pub admitted_s: f64,
    pub finished_s: f64,
    pub error: [c_char; 256],
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct hpt_server_stats {
    pub submitted: u64,
    pub finished: u64,
    pub failed: u64,
    pub bytes_in_use: u64,
    pub peak_bytes: u64,
    pub wavefront_steps: u64,
    pub max_jobs_per_step: u64,
    pub jobs_in_steps: u64,
    pub handles_reused: u64,
    pub cached_bytes: u64,
}

extern "C" {
    pub fn hpt_group_set_vmis_phases(g: *mut hpt_group, on: u32) -> hpt_status;
    pub fn hpt_server_config_init(c: *mut hpt_server_config);
    pub fn hpt_job_desc_init(d: *mut hpt_job_desc);
    pub fn hpt_server_create(config: *const hpt_server_config, out: *mut *mut hpt_server) -> hpt_status;
    pub fn hpt_server_add_scene(s: *mut hpt_server, scene: *const hpt_serve_scene_desc, scene_id: *mut u32)
        -> hpt_status;
    pub fn hpt_server_set_denoiser(s: *mut hpt_server, d: *mut hpt_denoiser) -> hpt_status;
    pub fn hpt_server_plan_job(s: *mut hpt_server, job: *const hpt_job_desc, footprint_bytes: *mut u64,
        chunk_frames: *mut u32) -> hpt_status;
    pub fn hpt_server_submit(s: *mut hpt_server, job: *const hpt_job_desc, job_id: *mut u64) -> hpt_status;
    pub fn hpt_server_wait(s: *mut hpt_server, job_id: u64, out: *mut hpt_job_result) -> hpt_status;
    pub fn hpt_server_next_result(s: *mut hpt_server, timeout_s: f64, out: *mut hpt_job_result, got: *mut c_int)
        -> hpt_status;
    pub fn hpt_server_read_image(s: *mut hpt_server, job_id: u64, dst: *mut f32, bytes: usize) -> hpt_status;
    pub fn hpt_server_drain(s: *mut hpt_server) -> hpt_status;
    pub fn hpt_server_clock(s: *mut hpt_server, seconds: *mut f64) -> hpt_status;
    pub fn hpt_server_get_stats(s: *mut hpt_server, out: *mut hpt_server_stats) -> hpt_status;
    pub fn hpt_server_destroy(s: *mut hpt_server);
}

/// The learned denoiser in place of denoise.py's oidnDenoise step
/// (denoise.py:26-30), run on the renderer's resident colour, albedo and
/// normal images.
pub struct Denoiser {
    raw: *mut hpt_denoiser,
}

impl Denoiser {
    /// `npz`: the bytes of the trained network's .npz file.
    pub fn new(npz: &[u8], device: i32) -> anyhow::Result<Self> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { hpt_denoiser_create(npz.as_ptr(), npz.len(), device, &mut raw) })?;
        Ok(Denoiser { raw })
    }

    /// Denoised colour, width*height RGBA f32 (alpha 1).
    pub fn denoise(&mut self, renderer: &mut HalaRenderer) -> anyhow::Result<Vec<f32>> {
        let mut v = vec![0f32; (renderer.width * renderer.height * 4) as usize];
        check(unsafe { hpt_denoise(self.raw, renderer.raw, v.as_mut_ptr(), v.len() * 4, 0) })?;
        Ok(v)
    }
}

impl Drop for Denoiser {
    fn drop(&mut self) {
        unsafe { hpt_denoiser_destroy(self.raw) }
    }
}

/// Opaque presenter handle (hpt_present_*).
#[repr(C)]
pub struct hpt_presenter {
    _private: [u8; 0],
}

/// The frame hand-off of the reference's window loop (update()/render(),
/// hala-pathtracer.rs:236-289): after each `render`, `present` queues a copy
/// of the final rgba8 image into a ring of pinned host buffers; `newest`
/// returns the latest finished one without waiting for the tracer, for the
/// window's texture upload.  Drop it before its renderer.
pub struct Presenter {
    raw: *mut hpt_presenter,
    width: u32,
    height: u32,
}

impl Presenter {
    pub fn new(renderer: &mut HalaRenderer, slots: u32) -> anyhow::Result<Self> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { hpt_present_create(renderer.raw, slots, &mut raw) })?;
        Ok(Presenter { raw, width: renderer.width, height: renderer.height })
    }

    pub fn present(&mut self) -> anyhow::Result<()> {
        check(unsafe { hpt_present_frame(self.raw) })
    }

    /// (frame count, rgba8 pixels) of the newest finished copy, valid until
    /// the next call; None before the first copy has finished.
    pub fn newest(&mut self) -> anyhow::Result<Option<(u64, &[u8])>> {
        let mut px: *const c_void = std::ptr::null();
        let mut frame = 0u64;
        check(unsafe { hpt_present_acquire(self.raw, &mut px, &mut frame) })?;
        if px.is_null() {
            return Ok(None);
        }
        let n = (self.width * self.height * 4) as usize;
        Ok(Some((frame, unsafe { std::slice::from_raw_parts(px as *const u8, n) })))
    }
}

impl Drop for Presenter {
    fn drop(&mut self) {
        unsafe { hpt_present_destroy(self.raw) }
    }
}

// The learned denoiser's kernels (include/hpt_nn.h), the replacement of the
// `oidnDenoise` step of denoise.py.  Streams are hipStream_t (null = the null
// stream); device pointers come from hpt_nn_alloc or hpt_get_image_pointer.
extern "C" {
    pub fn hpt_nn_conv_forward(stream: *mut c_void, x: *const u16, w: *const u16, b: *const f32, y: *mut u16,
        y_f32: *mut f32, n: c_int, h: c_int, wd: c_int, cin: c_int, cout: c_int, cout_pad: c_int, k: c_int,
        relu: c_int) -> c_int;
    pub fn hpt_nn_conv_forward_upcat(stream: *mut c_void, low: *const u16, skip: *const u16, w: *const u16,
        b: *const f32, y: *mut u16, y_f32: *mut f32, n: c_int, h: c_int, wd: c_int, c_low: c_int, c_skip: c_int,
        cout: c_int, cout_pad: c_int, relu: c_int) -> c_int;
    pub fn hpt_nn_conv_backward_data(stream: *mut c_void, dy: *const f32, wt: *const u16, dx: *mut f32, n: c_int,
        h: c_int, wd: c_int, cin: c_int, cout: c_int, cout_pad: c_int, k: c_int) -> c_int;
    pub fn hpt_nn_conv_backward_weights(stream: *mut c_void, dy: *const f32, x: *const u16, dw: *mut f32,
        db: *mut f32, partial: *mut f32, n: c_int, h: c_int, wd: c_int, cin: c_int, cout: c_int, cout_pad: c_int,
        k: c_int) -> c_int;
    pub fn hpt_nn_wgrad_scratch_floats(n: c_int, h: c_int, wd: c_int, cin: c_int, cout_pad: c_int, k: c_int) -> usize;
    pub fn hpt_nn_prepare_weights(stream: *mut c_void, master: *const f32, w: *mut u16, wt: *mut u16, cin: c_int,
        cout: c_int, cout_pad: c_int, k: c_int) -> c_int;
    pub fn hpt_nn_relu_backward(stream: *mut c_void, dy: *mut f32, y: *const u16, count: usize) -> c_int;
    pub fn hpt_nn_pool_forward(stream: *mut c_void, x: *const u16, y: *mut u16, n: c_int, h: c_int, wd: c_int,
        c: c_int) -> c_int;
    pub fn hpt_nn_pool_backward(stream: *mut c_void, dy: *const f32, dx: *mut f32, n: c_int, h: c_int, wd: c_int,
        c: c_int) -> c_int;
    pub fn hpt_nn_upsample_forward(stream: *mut c_void, x: *const u16, y: *mut u16, n: c_int, h: c_int, wd: c_int,
        c: c_int) -> c_int;
    pub fn hpt_nn_upsample_backward(stream: *mut c_void, dy: *const f32, dx: *mut f32, n: c_int, h: c_int,
        wd: c_int, c: c_int) -> c_int;
    pub fn hpt_nn_concat_forward(stream: *mut c_void, a: *const u16, b: *const u16, y: *mut u16, pixels: usize,
        ca: c_int, cb: c_int) -> c_int;
    pub fn hpt_nn_concat_backward(stream: *mut c_void, dy: *const f32, da: *mut f32, db: *mut f32, pixels: usize,
        ca: c_int, cb: c_int) -> c_int;
    pub fn hpt_nn_accumulate(stream: *mut c_void, dx: *mut f32, dy: *const f32, count: usize) -> c_int;
    pub fn hpt_nn_make_input(stream: *mut c_void, color: *const f32, albedo: *const f32, normal: *const f32,
        src_w: c_int, src_h: c_int, x0: c_int, y0: c_int, x: *mut u16, h: c_int, wd: c_int, mirror: c_int,
        aux: f32) -> c_int;
    pub fn hpt_nn_make_target(stream: *mut c_void, color: *const f32, src_w: c_int, src_h: c_int, x0: c_int,
        y0: c_int, out: *mut f32, h: c_int, wd: c_int, mirror: c_int) -> c_int;
    pub fn hpt_nn_add_residual(stream: *mut c_void, out: *mut f32, x: *const u16, pixels: usize,
        x_channels: c_int) -> c_int;
    pub fn hpt_nn_l1_loss(stream: *mut c_void, out: *const f32, target_log: *const f32, dy: *mut f32,
        loss_sum: *mut f32, pixels: usize, cout_pad: c_int) -> c_int;
    pub fn hpt_nn_adam(stream: *mut c_void, p: *mut f32, g: *const f32, m: *mut f32, v: *mut f32, count: usize,
        lr: f32, beta1: f32, beta2: f32, eps: f32, t: c_int) -> c_int;
    pub fn hpt_nn_write_output(stream: *mut c_void, out: *const f32, dst: *mut f32, dst_w: c_int, dst_h: c_int,
        x0: c_int, y0: c_int, h: c_int, wd: c_int) -> c_int;
    pub fn hpt_nn_alloc(bytes: usize, ptr: *mut *mut c_void) -> c_int;
    pub fn hpt_nn_free(ptr: *mut c_void) -> c_int;
    pub fn hpt_nn_copy(stream: *mut c_void, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn hpt_nn_fill(stream: *mut c_void, p: *mut f32, value: f32, count: usize) -> c_int;
    pub fn hpt_nn_stream_synchronize(stream: *mut c_void) -> c_int;
    /// A non-blocking stream for the denoiser's work (never the null stream).
    pub fn hpt_nn_stream_create(stream: *mut *mut c_void) -> c_int;
    pub fn hpt_nn_event_create(event: *mut *mut c_void) -> c_int;
    pub fn hpt_nn_event_destroy(event: *mut c_void) -> c_int;
    pub fn hpt_nn_event_record(event: *mut c_void, stream: *mut c_void) -> c_int;
    pub fn hpt_nn_event_elapsed(start: *mut c_void, stop: *mut c_void, ms: *mut f32) -> c_int;
}

fn check(status: hpt_status) -> anyhow::Result<()> {
    if status == 0 {
        return Ok(());
    }
    let msg = unsafe { CStr::from_ptr(hpt_last_error()) }.to_string_lossy().into_owned();
    Err(anyhow::anyhow!("hpt status {}: {}", status, msg))
}

/// Owned mesh data of one primitive (what `HalaScene` holds per primitive).
pub struct Mesh {
    pub transform: [f32; 16],
    pub material_index: u32,
    pub vertices: Vec<hpt_vertex>,
    pub indices: Vec<u32>,
}

/// The scene arrays the renderer uploads in `set_scene`.
pub struct SceneArrays {
    pub cameras: Vec<hpt_camera>,
    pub lights: Vec<hpt_light>,
    pub materials: Vec<hpt_material>,
    pub meshes: Vec<Mesh>,
}

pub struct HalaRenderer {
    raw: *mut hpt_renderer,
    width: u32,
    height: u32,
}

impl HalaRenderer {
    /// hala-pathtracer.rs:128-138, without the window; `features` is the
    /// permutation the host picks at :101-114.
    #[allow(clippy::too_many_arguments)]
    pub fn new(width: u32, height: u32, max_depth: u32, rr_depth: u32, enable_tonemap: bool, enable_aces: bool,
               use_simple_aces: bool, max_samples: u64, features: u32) -> anyhow::Result<Self> {
        let info = hpt_create_info {
            width,
            height,
            max_depth,
            rr_depth,
            enable_tonemap: enable_tonemap as u32,
            enable_aces: enable_aces as u32,
            use_simple_aces: use_simple_aces as u32,
            features,
            max_samples,
            exposure_value: 0.0,
            device: 0,
        };
        let mut raw = std::ptr::null_mut();
        check(unsafe { hpt_create(&info, &mut raw) })?;
        Ok(Self { raw, width, height })
    }

    /// hala-pathtracer.rs:198.
    pub fn set_scene(&mut self, scene: &SceneArrays) -> anyhow::Result<()> {
        let prims: Vec<hpt_primitive_desc> = scene
            .meshes
            .iter()
            .map(|m| hpt_primitive_desc {
                transform: m.transform,
                material_index: m.material_index,
                vertex_count: m.vertices.len() as u32,
                index_count: m.indices.len() as u32,
                _pad: 0,
                vertices: m.vertices.as_ptr(),
                indices: m.indices.as_ptr(),
            })
            .collect();
        let desc = hpt_scene_desc {
            cameras: scene.cameras.as_ptr(),
            camera_count: scene.cameras.len() as u32,
            light_count: scene.lights.len() as u32,
            lights: scene.lights.as_ptr(),
            materials: scene.materials.as_ptr(),
            material_count: scene.materials.len() as u32,
            primitive_count: prims.len() as u32,
            primitives: prims.as_ptr(),
        };
        check(unsafe { hpt_set_scene(self.raw, &desc) })
    }

    /// hala-pathtracer.rs:203 after the HDR file is decoded to linear RGB.
    pub fn set_envmap(&mut self, rgb: &[f32], width: u32, height: u32, rotation: f32) -> anyhow::Result<()> {
        anyhow::ensure!(rgb.len() == (width * height * 3) as usize, "envmap size mismatch");
        check(unsafe { hpt_set_env_map(self.raw, rgb.as_ptr(), width, height, rotation) })
    }

    /// hala-pathtracer.rs:203 as the reference calls it, with an OpenEXR path.
    pub fn set_envmap_exr(&mut self, path: &str, rotation: f32) -> anyhow::Result<()> {
        let bytes = std::fs::read(path)?;
        let (mut width, mut height) = (0u32, 0u32);
        check(unsafe { hpt_exr_info(bytes.as_ptr(), bytes.len(), &mut width, &mut height) })?;
        let mut rgb = vec![0f32; (width as usize) * (height as usize) * 3];
        check(unsafe { hpt_exr_decode_rgb(bytes.as_ptr(), bytes.len(), rgb.as_mut_ptr(), width, height) })?;
        self.set_envmap(&rgb, width, height, rotation)
    }

    pub fn set_ground_color(&mut self, c: [f32; 4]) -> anyhow::Result<()> {
        check(unsafe { hpt_set_ground_color(self.raw, c.as_ptr()) })
    }

    pub fn set_sky_color(&mut self, c: [f32; 4]) -> anyhow::Result<()> {
        check(unsafe { hpt_set_sky_color(self.raw, c.as_ptr()) })
    }

    pub fn set_env_intensity(&mut self, v: f32) -> anyhow::Result<()> {
        check(unsafe { hpt_set_env_intensity(self.raw, v) })
    }

    pub fn set_exposure_value(&mut self, ev: f32) -> anyhow::Result<()> {
        check(unsafe { hpt_set_exposure_value(self.raw, ev) })
    }

    /// hala-pathtracer.rs:195 after the PNG is decoded to RGBA8.
    pub fn load_blue_noise_texture(&mut self, rgba8: &[u8], width: u32, height: u32) -> anyhow::Result<()> {
        check(unsafe { hpt_set_blue_noise(self.raw, rgba8.as_ptr(), width, height) })
    }

    /// hala-pathtracer.rs:213.
    /// The g_textures[] table of the scene's materials: (width, height, RGBA8 texels).
    pub fn set_textures(&mut self, textures: &[(u32, u32, &[u8])]) -> anyhow::Result<()> {
        for (w, h, t) in textures {
            anyhow::ensure!(t.len() == (*w as usize) * (*h as usize) * 4, "texture size mismatch");
        }
        let descs: Vec<hpt_texture_desc> = textures
            .iter()
            .map(|(w, h, t)| hpt_texture_desc { width: *w, height: *h, rgba8: t.as_ptr() })
            .collect();
        check(unsafe { hpt_set_textures(self.raw, descs.as_ptr(), descs.len() as u32) })
    }

    pub fn commit(&mut self) -> anyhow::Result<()> {
        check(unsafe { hpt_commit(self.raw) })
    }

    /// The update()+render() pair of hala-pathtracer.rs:264-285: one frame.
    pub fn render(&mut self, camera_index: u32) -> anyhow::Result<()> {
        check(unsafe { hpt_render(self.raw, 1, camera_index) })
    }

    pub fn reset_accumulation(&mut self) -> anyhow::Result<()> {
        check(unsafe { hpt_reset_accumulation(self.raw) })
    }

    /// Linear RGBA32F planes behind save_images (hala-pathtracer.rs:251).
    pub fn read_rgba32f(&self, which: c_int) -> anyhow::Result<Vec<f32>> {
        let mut v = vec![0f32; (self.width * self.height * 4) as usize];
        check(unsafe { hpt_read_image(self.raw, which, v.as_mut_ptr() as *mut c_void, v.len() * 4) })?;
        Ok(v)
    }

    pub fn read_final_rgba8(&self) -> anyhow::Result<Vec<u8>> {
        let mut v = vec![0u8; (self.width * self.height * 4) as usize];
        check(unsafe { hpt_read_image(self.raw, HPT_IMAGE_FINAL_RGBA8, v.as_mut_ptr() as *mut c_void, v.len()) })?;
        Ok(v)
    }

    /// hala-pathtracer.rs:228.
    pub fn wait_idle(&self) -> anyhow::Result<()> {
        check(unsafe { hpt_wait_idle(self.raw) })
    }
}

impl Drop for HalaRenderer {
    fn drop(&mut self) {
        unsafe { hpt_destroy(self.raw) }
    }
}
