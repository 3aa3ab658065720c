/*
 * camera.cpp - CameraParameters / CameraController restated from
 * src/camera/cameraParameters.cpp:7-186 and src/camera/cameraController.cpp:4-24
 * (same defaults, same float formulas, same clamping).
 */
#include "camera.h"

#include <cmath>

namespace voxeltoy {

/* cameraParameters.cpp:7-17: target origin, eye at -Z, near 0.1, far 10000,
 * focal distance 100, pinhole, 36x36 film, 50 mm focal length. */
CameraParameters::CameraParameters()
    : m_target(0.0f, 0.0f, 0.0f), m_eye(0.0f, 0.0f, -1.0f), m_near(0.1f), m_far(10000.0f),
      m_focalDistance(100.0f), m_lensRadius(0.0f), m_filmSize(36.0f), m_lensModel(CLM_PINHOLE)
{
    setFocalLength(50.0f);
}

float CameraParameters::distanceToTarget() const { return (m_target - m_eye).length(); }

void CameraParameters::lookAt(const V3f& target) { m_target = target; }

void CameraParameters::setDistanceFromTarget(float distance)
{
    m_eye = m_target - forwardUnitVector() * distance;
}

void CameraParameters::getBasis(V3f& forward, V3f& right, V3f& up) const
{
    forward = (m_target - m_eye).normalized();
    right = V3f(0.0f, 1.0f, 0.0f).cross(forward).normalized();
    up = forward.cross(right);
}

V3f CameraParameters::forwardUnitVector() const
{
    V3f f, r, u;
    getBasis(f, r, u);
    return f;
}
V3f CameraParameters::rightUnitVector() const
{
    V3f f, r, u;
    getBasis(f, r, u);
    return r;
}
V3f CameraParameters::upUnitVector() const
{
    V3f f, r, u;
    getBasis(f, r, u);
    return u;
}

float CameraParameters::rotationTheta() const { return std::acos(forwardUnitVector().y); }
float CameraParameters::rotationPhi() const
{
    const V3f f = forwardUnitVector();
    return std::atan2(f.z, f.x);
}

static V3f spherical_to_cartesian(float theta, float phi)
{
    const float st = std::sin(theta);
    return V3f(st * std::cos(phi), std::cos(theta), st * std::sin(phi));
}

void CameraParameters::orbitAroundTarget(float theta, float phi)
{
    const float r = distanceToTarget();
    m_eye = m_target - r * spherical_to_cartesian(theta, phi);
}

void CameraParameters::orbitAroundEye(float theta, float phi)
{
    const float r = distanceToTarget();
    m_target = m_eye + r * spherical_to_cartesian(theta, phi);
}

void CameraParameters::setEyeTarget(const V3f& eye, const V3f& target)
{
    m_eye = eye;
    m_target = target;
}

void CameraParameters::setFovY(float fov) { m_fovY = fov; }
void CameraParameters::setNearDistance(float d) { m_near = std::max(0.0f, d); }
void CameraParameters::setFarDistance(float d) { m_far = std::max(m_near, d); }

float CameraParameters::focalLength() const { return filmSize().y / (2.0f * std::tan(0.5f * m_fovY)); }

void CameraParameters::setFocalLength(float focalLength)
{
    m_fovY = std::atan2(filmSize().y * 0.5f, focalLength) * 2.0f;
}

void CameraParameters::setFocalDistance(float distance) { m_focalDistance = std::max(0.0f, distance); }

/* The focal length is re-evaluated after the film changes, so the fov is
 * kept (up to rounding), as cameraParameters.cpp:157-162 does. */
void CameraParameters::setFilmSize(float filmW, float filmH)
{
    m_filmSize = V2f(std::max(0.0f, filmW), std::max(0.0f, filmH));
    m_fovY = std::atan2(filmSize().y * 0.5f, focalLength()) * 2.0f;
}

void CameraParameters::setLensRadius(float radius) { m_lensRadius = std::max(0.0f, radius); }

void CameraParameters::setFStop(float fstop) { m_lensRadius = (focalLength() / std::max(1e-4f, fstop)) * 0.5f; }

void CameraParameters::setLensModel(CameraLensModel model) { m_lensModel = model; }

void CameraController::lookAt(const V3f& target)
{
    if (m_parameters) m_parameters->lookAt(target);
}

void CameraController::setDistanceFromTarget(float distance)
{
    if (m_parameters) m_parameters->setDistanceFromTarget(distance);
}

void CameraController::focusOnBounds(const Box3f& bounds)
{
    if (!m_parameters) return;
    const V3f fwd = m_parameters->forwardUnitVector();
    const V3f center = bounds.center();
    const float distance = bounds.size().length() / (2.0f * std::tan(m_parameters->fovY() / 2));
    m_parameters->setEyeTarget(center - fwd * distance, center);
}

void CameraController::orbit(float theta, float phi)
{
    if (m_parameters) m_parameters->orbitAroundTarget(theta, phi);
}

void CameraController::look(float theta, float phi)
{
    if (m_parameters) m_parameters->orbitAroundEye(theta, phi);
}

void CameraController::setEyeTarget(const V3f& eye, const V3f& target)
{
    if (m_parameters) m_parameters->setEyeTarget(eye, target);
}

/* The angle arithmetic is the controllers' own: the float offsets are
 * promoted to double by M_PI and the sum rounded to float once
 * (orbitCameraController.cpp:20-24, flyCameraController.cpp:20-24). */
bool CameraController::onMouseMove(float dx, float dy, int buttons)
{
    if (!m_parameters) return false;
    const float speed = 1.0f;
    if (buttons & 2) {
        const float theta = (float)((double)dy * M_PI * speed + m_parameters->rotationTheta());
        const float phi = (float)(-(double)dx * 2.0 * M_PI * speed + m_parameters->rotationPhi());
        if (m_mode == FLY)
            m_parameters->orbitAroundEye(theta, phi);
        else
            m_parameters->orbitAroundTarget(theta, phi);
        return true;
    }
    if (buttons & 4) {
        if (m_mode == FLY) {
            const V3f up = dy * m_parameters->upUnitVector() * 100.0f;
            m_parameters->setEyeTarget(m_parameters->eye() + up, m_parameters->target() + up);
        } else {
            const float k = 1.05f;
            const float d = m_parameters->distanceToTarget();
            m_parameters->setDistanceFromTarget(dy > 0 ? d * k : d / k);
        }
        return true;
    }
    return false;
}

bool CameraController::onKeyPress(int key)
{
    if (!m_parameters || m_mode != FLY) return false;
    const float step = 100.0f;
    V3f move;
    switch (key) {
    case 'W': move = m_parameters->forwardUnitVector() * step; break;
    case 'S': move = -(m_parameters->forwardUnitVector() * step); break;
    case 'D': move = m_parameters->rightUnitVector() * step; break;
    case 'A': move = -(m_parameters->rightUnitVector() * step); break;
    default: return false;
    }
    m_parameters->setEyeTarget(m_parameters->eye() + move, m_parameters->target() + move);
    return true;
}

}  // namespace voxeltoy
