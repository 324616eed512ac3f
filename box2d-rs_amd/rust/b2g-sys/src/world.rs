//! Safe wrapper over one GPU world, keeping `B2world::step`'s signature (src/b2_world.rs:98-100).
//! In box2d-rs it sits behind a cargo feature (`gpu`): `B2world::new_gpu(gravity)` owns a
//! `GpuWorld`, bodies and fixtures created through the usual API are staged here and handed to the
//! device in creation order before the next step (so proxy ids and list orders match the
//! reference), and `B2body` getters read the states synchronised after a step. The world is owned
//! by one Rust object and used from one thread, as the reference's `Rc<RefCell<…>>` world is.
use super::*;
use std::ffi::CStr;

/// A b2g status other than B2G_OK, with the library's message.
#[derive(Debug, Clone)]
pub struct Error {
    pub status: i32,
    pub message: String,
}

fn check(status: c_int) -> Result<(), Error> {
    if status == B2G_OK {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(b2g_last_error()) }.to_string_lossy().into_owned();
    Err(Error { status, message })
}

pub struct GpuWorld {
    raw: *mut b2g_world,
    staged_bodies: Vec<b2g_body_def>,
    staged_fixture_body: Vec<i32>,
    staged_shapes: Vec<b2g_shape_def>,
    staged_fixtures: Vec<b2g_fixture_def>,
    body_count: i32,
    states: Vec<b2g_body_state>,
}

impl GpuWorld {
    /// B2world::new(gravity): warm starting, sleeping and continuous physics on.
    pub fn new(gravity_x: f32, gravity_y: f32) -> Result<Self, Error> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { b2g_world_create(gravity_x, gravity_y, &mut raw) })?;
        Ok(GpuWorld {
            raw,
            staged_bodies: Vec::new(),
            staged_fixture_body: Vec::new(),
            staged_shapes: Vec::new(),
            staged_fixtures: Vec::new(),
            body_count: 0,
            states: Vec::new(),
        })
    }

    /// B2world::set_warm_starting / set_allow_sleeping / set_continuous_physics in one call.
    pub fn set_flags(&mut self, warm_starting: bool, allow_sleep: bool, continuous: bool) -> Result<(), Error> {
        self.flush()?;
        check(unsafe { b2g_world_set_flags(self.raw, warm_starting as i32, allow_sleep as i32, continuous as i32, 1) })
    }

    /// B2world::create_body: returns the body's index (creation order).
    pub fn create_body(&mut self, def: b2g_body_def) -> i32 {
        self.staged_bodies.push(def);
        self.body_count + self.staged_bodies.len() as i32 - 1
    }

    /// B2body::create_fixture on a body returned by create_body.
    pub fn create_fixture(&mut self, body: i32, shape: b2g_shape_def, fixture: b2g_fixture_def) {
        self.staged_fixture_body.push(body);
        self.staged_shapes.push(shape);
        self.staged_fixtures.push(fixture);
    }

    /// Hands the staged bodies and fixtures to the device in creation order.
    fn flush(&mut self) -> Result<(), Error> {
        if self.staged_bodies.is_empty() && self.staged_fixtures.is_empty() {
            return Ok(());
        }
        let base = self.body_count;
        let fixture_body: Vec<i32> = self.staged_fixture_body.iter().map(|b| b - base).collect();
        check(unsafe {
            b2g_world_add(self.raw, self.staged_bodies.len() as i32, self.staged_bodies.as_ptr(),
                          self.staged_fixtures.len() as i32, fixture_body.as_ptr(), self.staged_shapes.as_ptr(),
                          self.staged_fixtures.as_ptr(), std::ptr::null_mut())
        })?;
        self.body_count += self.staged_bodies.len() as i32;
        self.staged_bodies.clear();
        self.staged_fixture_body.clear();
        self.staged_shapes.clear();
        self.staged_fixtures.clear();
        Ok(())
    }

    /// B2world::step(dt, velocity_iterations, position_iterations); the body states are read back
    /// afterwards for the getters.
    pub fn step(&mut self, dt: f32, velocity_iterations: i32, position_iterations: i32) -> Result<(), Error> {
        self.flush()?;
        check(unsafe { b2g_world_step(self.raw, dt, velocity_iterations, position_iterations) })?;
        self.states.resize(self.body_count as usize, b2g_body_state::default());
        check(unsafe { b2g_world_get_body_states(self.raw, self.states.as_mut_ptr(), self.body_count) })
    }

    /// B2body::get_position, get_angle, get_linear_velocity, get_angular_velocity after the last step.
    pub fn body_state(&self, body: i32) -> Option<&b2g_body_state> {
        self.states.get(body as usize)
    }

    /// B2world::get_profile.
    pub fn profile(&self) -> Result<b2g_profile, Error> {
        let mut p = b2g_profile::default();
        check(unsafe { b2g_world_get_profile(self.raw, &mut p) })?;
        Ok(p)
    }
}

impl Drop for GpuWorld {
    fn drop(&mut self) {
        unsafe {
            b2g_world_destroy(self.raw);
        }
    }
}
